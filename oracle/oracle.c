/*
 * oracle.c -- CPU restatement of the reference's 2-D ideal-MHD update.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests compare the HIP product path with
 * this library, and it is the CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in the product (aeromhd-cuda_amd/) links or calls it.
 *
 * It follows the reference (Semedo2002/AeroMHD-CUDA) function by function,
 * on the reference's own ghosted layout: every state array is
 * (NVAR, nxt, nyt) in C order with nxt = nx + 4, nyt = ny + 4 (ng = 2),
 * y fastest.  Every arithmetic expression keeps the reference's operation
 * order so that an FP64 build with -ffp-contract=off reproduces IEEE
 * per-operation results (no FMA contraction, true division, IEEE sqrt).
 *
 * Function map (reference file:line):
 *   orc_cons_to_prim        physics.py:284-301
 *   orc_prim_to_cons        physics.py:304-320
 *   orc_muscl_x / _y        physics.py:323-350
 *   orc_hlld_flux           physics.py:43-275 (kernel), 353-379 (wrappers)
 *   orc_apply_bc            solver.py:158-197
 *   orc_enforce_bounds      solver.py:199-204
 *   orc_compute_dt          solver.py:206-223
 *   orc_boundary_flux       solver.py:225-240
 *   orc_powell_source       solver.py:242-266
 *   orc_compute_rhs         solver.py:268-308
 *   orc_step_ssprk3         solver.py:310-333
 *   orc_run                 solver.py:475-482 / main.py:66-73 time loop
 *
 * Pinning (what checks each function independently of this text) is listed
 * in the header of oracle/oracle.py and in the tests under tests/.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NV 10
/* conserved order  [rho, mx, my, mz, Bx, By, Bz, E, psi, rhoC]  (config.py:13-16)
 * primitive order  [rho, vx, vy, vz, Bx, By, Bz, p, psi, C]     (config.py:19-22) */
enum { Q_RHO = 0, Q_MX, Q_MY, Q_MZ, Q_BX, Q_BY, Q_BZ, Q_EN, Q_PSI, Q_RHOC };
enum { P_RHO = 0, P_VX, P_VY, P_VZ, P_BX, P_BY, P_BZ, P_PR, P_PSI, P_CLR };
static const double FLOOR_RHO = 1e-12; /* config.py:40 */
static const double FLOOR_PR = 1e-12;  /* config.py:41 */
#define NG 2

typedef struct {
    int nx, ny;          /* interior cells */
    double gamma, dx, dy;
    int bcx_periodic;    /* 1: periodic, 0: extrapolation / characteristic */
    int bcy_periodic;
    int powell;          /* powell_source && B_transverse > 0 (solver.py:304) */
    double glm_alpha;
} orc_params;

static inline double dmax(double a, double b) { return a > b ? a : b; }

#ifdef _OPENMP
#include <omp.h>
#endif

/* Thread count for the CPU baseline (1 = scalar). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- */
/* physics.py:284-301 for one cell: u, w are NV-vectors.              */
static inline void c2p_cell(const double u[NV], double w[NV], double gamma)
{
    double rho = dmax(u[Q_RHO], FLOOR_RHO);
    double ri = 1.0 / rho;
    double vx = u[Q_MX] * ri, vy = u[Q_MY] * ri, vz = u[Q_MZ] * ri;
    double bx = u[Q_BX], by = u[Q_BY], bz = u[Q_BZ];
    w[P_RHO] = rho; w[P_VX] = vx; w[P_VY] = vy; w[P_VZ] = vz;
    w[P_BX] = bx; w[P_BY] = by; w[P_BZ] = bz;
    w[P_PSI] = u[Q_PSI];
    w[P_CLR] = u[Q_RHOC] * ri;
    /* KE = 0.5*rho*(vx^2+vy^2+vz^2); ME = 0.5*(Bx^2+By^2+Bz^2);
       p = max((gamma-1)*(E - KE - ME), floor), sums left to right */
    double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    double me = 0.5 * (bx * bx + by * by + bz * bz);
    w[P_PR] = dmax((gamma - 1) * (u[Q_EN] - ke - me), FLOOR_PR);
}

/* physics.py:284-301.  U, W: (NV, n) planes. */
void orc_cons_to_prim(const double* U, double* W, long n, double gamma)
{
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        double u[NV], w[NV];
        for (int v = 0; v < NV; ++v) u[v] = U[v * n + k];
        c2p_cell(u, w, gamma);
        for (int v = 0; v < NV; ++v) W[v * n + k] = w[v];
    }
}

/* physics.py:304-320 */
void orc_prim_to_cons(const double* W, double* U, long n, double gamma)
{
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        double rho = dmax(W[P_RHO * n + k], FLOOR_RHO);
        double vx = W[P_VX * n + k], vy = W[P_VY * n + k], vz = W[P_VZ * n + k];
        double bx = W[P_BX * n + k], by = W[P_BY * n + k], bz = W[P_BZ * n + k];
        U[Q_RHO * n + k] = rho;
        U[Q_MX * n + k] = rho * vx;
        U[Q_MY * n + k] = rho * vy;
        U[Q_MZ * n + k] = rho * vz;
        U[Q_BX * n + k] = bx;
        U[Q_BY * n + k] = by;
        U[Q_BZ * n + k] = bz;
        U[Q_PSI * n + k] = W[P_PSI * n + k];
        U[Q_RHOC * n + k] = rho * W[P_CLR * n + k];
        double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
        double me = 0.5 * (bx * bx + by * by + bz * bz);
        U[Q_EN * n + k] = W[P_PR * n + k] / (gamma - 1) + ke + me;
    }
}

/* van Leer slope of physics.py:325-328: where(ab > 0, 2*ab/(dm+dp+1e-30), 0) */
static inline double vl_slope(double wm, double w0, double wp)
{
    double dm = w0 - wm, dp = wp - w0;
    double ab = dm * dp;
    return ab > 0 ? (2.0 * ab) / (dm + dp + 1e-30) : 0.0;
}

/* physics.py:323-335.  W (NV, nxt, nyt) -> WL, WR (NV, nxt-3, nyt).
 * Face k (0..nxt-4) lies between ghosted cells k+1 and k+2. */
void orc_muscl_x(const double* W, int nxt, int nyt, double* WL, double* WR)
{
    long plane = (long)nxt * nyt, fplane = (long)(nxt - 3) * nyt;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nxt - 3; ++k) {
        for (int v = 0; v < NV; ++v) {
            const double* w = W + v * plane;
            for (int j = 0; j < nyt; ++j) {
                double d1 = vl_slope(w[(long)k * nyt + j], w[(long)(k + 1) * nyt + j], w[(long)(k + 2) * nyt + j]);
                double d2 = vl_slope(w[(long)(k + 1) * nyt + j], w[(long)(k + 2) * nyt + j], w[(long)(k + 3) * nyt + j]);
                double l = w[(long)(k + 1) * nyt + j] + 0.5 * d1;
                double r = w[(long)(k + 2) * nyt + j] - 0.5 * d2;
                if (v == P_RHO) { l = dmax(l, FLOOR_RHO); r = dmax(r, FLOOR_RHO); }
                if (v == P_PR)  { l = dmax(l, FLOOR_PR);  r = dmax(r, FLOOR_PR); }
                WL[v * fplane + (long)k * nyt + j] = l;
                WR[v * fplane + (long)k * nyt + j] = r;
            }
        }
    }
}

/* physics.py:338-350.  W (NV, nxt, nyt) -> WL, WR (NV, nxt, nyt-3). */
void orc_muscl_y(const double* W, int nxt, int nyt, double* WL, double* WR)
{
    long plane = (long)nxt * nyt;
    int nf = nyt - 3;
    long fplane = (long)nxt * nf;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nxt; ++i) {
        for (int v = 0; v < NV; ++v) {
            const double* w = W + v * plane + (long)i * nyt;
            for (int k = 0; k < nf; ++k) {
                double d1 = vl_slope(w[k], w[k + 1], w[k + 2]);
                double d2 = vl_slope(w[k + 1], w[k + 2], w[k + 3]);
                double l = w[k + 1] + 0.5 * d1;
                double r = w[k + 2] - 0.5 * d2;
                if (v == P_RHO) { l = dmax(l, FLOOR_RHO); r = dmax(r, FLOOR_RHO); }
                if (v == P_PR)  { l = dmax(l, FLOOR_PR);  r = dmax(r, FLOOR_PR); }
                WL[v * fplane + (long)i * nf + k] = l;
                WR[v * fplane + (long)i * nf + k] = r;
            }
        }
    }
}

/* ---------------------------------------------------------------- */
/* HLLD face solver, physics.py:43-275.                              */

/* One side of a face after the common GLM / wave-speed set-up. */
typedef struct {
    double rho, vx, vy, vz, by, bz, p, psi, c;
    double bm2, pt, e, vb;       /* |B|^2 with face Bx, total pressure, energy, v.B */
    double ds;                   /* S_K - vx_K */
    double ds_sm;                /* S_K - SM (clamped, sign discarded) */
    double rhos, sq;             /* star density, sqrt(star density) */
    double vys, vzs, bys, bzs, es, vbs;   /* star state */
    double eds;                  /* double-star energy */
} orc_side;

/* Physical x-flux of one side (physics.py:178-187 / 189-198 / 202-211). */
static void side_flux(const orc_side* s, double bxf, double bx2, double psif, double ch, double f[NV])
{
    f[0] = s->rho * s->vx;
    f[1] = s->rho * s->vx * s->vx + s->pt - bx2;
    f[2] = s->rho * s->vy * s->vx - bxf * s->by;
    f[3] = s->rho * s->vz * s->vx - bxf * s->bz;
    f[4] = (ch > 0.0) ? psif : 0.0;
    f[5] = s->by * s->vx - bxf * s->vy;
    f[6] = s->bz * s->vx - bxf * s->vz;
    f[7] = (s->e + s->pt) * s->vx - bxf * s->vb;
    f[8] = (ch > 0.0) ? bxf * ch * ch : 0.0;
    f[9] = s->rho * s->c * s->vx;
}

/* Conserved vectors of one side: original, star, double star (physics.py:213-249). */
static void side_states(const orc_side* s, double bxf, double psif, double sm,
                        double vyds, double vzds, double byds, double bzds,
                        double u[NV], double us[NV], double uds[NV])
{
    u[0] = s->rho; u[1] = s->rho * s->vx; u[2] = s->rho * s->vy; u[3] = s->rho * s->vz;
    u[4] = bxf; u[5] = s->by; u[6] = s->bz; u[7] = s->e; u[8] = s->psi; u[9] = s->rho * s->c;

    us[0] = s->rhos; us[1] = s->rhos * sm; us[2] = s->rhos * s->vys; us[3] = s->rhos * s->vzs;
    us[4] = bxf; us[5] = s->bys; us[6] = s->bzs; us[7] = s->es; us[8] = psif; us[9] = s->rhos * s->c;

    uds[0] = s->rhos; uds[1] = s->rhos * sm; uds[2] = s->rhos * vyds; uds[3] = s->rhos * vzds;
    uds[4] = bxf; uds[5] = byds; uds[6] = bzds; uds[7] = s->eds; uds[8] = psif; uds[9] = s->rhos * s->c;
}

static void load_side(orc_side* s, const double w[NV], double floor_rho, double floor_pr)
{
    s->rho = w[0]; if (s->rho < floor_rho) s->rho = floor_rho;
    s->vx = w[1]; s->vy = w[2]; s->vz = w[3];
    s->by = w[5]; s->bz = w[6];
    s->p = w[7]; if (s->p < floor_pr) s->p = floor_pr;
    s->psi = w[8]; s->c = w[9];
}

/* Fast magnetosonic speed, physics.py:94-97 */
static double fast_speed(const orc_side* s, double gamma, double bx2)
{
    double a2 = gamma * s->p / s->rho;
    double tmp = a2 + s->bm2 / s->rho;
    double disc = tmp * tmp - 4.0 * a2 * bx2 / s->rho;
    if (disc < 0.0) disc = 0.0;
    return sqrt(fmax(0.5 * (tmp + sqrt(disc)), 0.0));
}

/* Star state of one side, physics.py:122-154.  sgn_ds_sm: the clamp keeps
 * |S_K - SM| >= 1e-14 and discards the sign (physics.py:122-123). */
static void star_side(orc_side* s, double sk, double sm, double pts, double bxf, double bx2,
                      double gm1, double floor_rho)
{
    s->ds_sm = sk - sm; if (fabs(s->ds_sm) < 1e-14) s->ds_sm = 1e-14;
    s->rhos = fmax(s->rho * s->ds / s->ds_sm, floor_rho);
    s->sq = sqrt(s->rhos);
    double eps_rel = 1e-8 * fmax(fabs(pts), 1.0);
    double dn = s->rho * s->ds * s->ds_sm - bx2;
    if (fabs(dn) < eps_rel) dn = (dn >= 0.0) ? eps_rel : -eps_rel;
    s->vys = s->vy - bxf * s->by * (sm - s->vx) / dn;
    s->vzs = s->vz - bxf * s->bz * (sm - s->vx) / dn;
    s->bys = s->by * (s->rho * s->ds * s->ds - bx2) / dn;
    s->bzs = s->bz * (s->rho * s->ds * s->ds - bx2) / dn;
    s->vb = s->vx * bxf + s->vy * s->by + s->vz * s->bz;
    s->vbs = sm * bxf + s->vys * s->bys + s->vzs * s->bzs;
    s->e = s->p / gm1 + 0.5 * s->rho * (s->vx * s->vx + s->vy * s->vy + s->vz * s->vz) + 0.5 * s->bm2;
    s->es = ((sk - s->vx) * s->e - s->pt * s->vx + pts * sm + bxf * (s->vb - s->vbs)) / s->ds_sm;
}

/* One x-face.  wl, wr: primitive vectors (already rotated for y faces).
 * Returns smax = max(|SL|, |SR|). */
static double hlld_face(const double wl[NV], const double wr[NV], double f[NV],
                        double gamma, double ch, double floor_rho, double floor_pr)
{
    orc_side L, R;
    double gm1 = gamma - 1.0;
    load_side(&L, wl, floor_rho, floor_pr);
    load_side(&R, wr, floor_rho, floor_pr);
    double bxl = wl[4], bxr = wr[4];

    /* GLM face solution, physics.py:74-82 */
    double bxf, psif;
    if (ch > 0.0) {
        bxf = 0.5 * (bxl + bxr) - 0.5 / ch * (R.psi - L.psi);
        psif = 0.5 * (L.psi + R.psi) - 0.5 * ch * (bxr - bxl);
    } else {
        bxf = 0.5 * (bxl + bxr);
        psif = 0.0;
    }
    double bx2 = bxf * bxf;
    L.bm2 = bx2 + L.by * L.by + L.bz * L.bz;
    R.bm2 = bx2 + R.by * R.by + R.bz * R.bz;
    L.pt = L.p + 0.5 * L.bm2;
    R.pt = R.p + 0.5 * R.bm2;

    /* wave speeds, physics.py:93-109 */
    double cfl = fast_speed(&L, gamma, bx2);
    double cfr = fast_speed(&R, gamma, bx2);
    double sl = fmin(L.vx - cfl, R.vx - cfr);
    double sr = fmax(L.vx + cfl, R.vx + cfr);
    if (ch > 0.0) { sl = fmin(sl, -ch); sr = fmax(sr, ch); }
    double smax = fmax(fabs(sl), fabs(sr));

    /* contact speed and star total pressure, physics.py:111-120 */
    L.ds = sl - L.vx;
    R.ds = sr - R.vx;
    double den = R.ds * R.rho - L.ds * L.rho;
    if (fabs(den) < 1e-14) den = (den >= 0.0) ? 1e-14 : -1e-14;
    double sm = (R.ds * R.rho * R.vx - L.ds * L.rho * L.vx - R.pt + L.pt) / den;
    if (sm < sl + 1e-14) sm = sl + 1e-14;
    if (sm > sr - 1e-14) sm = sr - 1e-14;
    double pts = L.pt + L.rho * L.ds * (sm - L.vx);

    star_side(&L, sl, sm, pts, bxf, bx2, gm1, floor_rho);
    star_side(&R, sr, sm, pts, bxf, bx2, gm1, floor_rho);

    /* double-star state, physics.py:156-172 */
    double sbx = (bxf >= 0.0) ? 1.0 : -1.0;
    double sql = fmax(L.sq, 1e-14), sqr = fmax(R.sq, 1e-14);
    double sal = sm - fabs(bxf) / sql;
    double sar = sm + fabs(bxf) / sqr;
    if (sal < sl) sal = sl;
    if (sal > sm) sal = sm;
    if (sar < sm) sar = sm;
    if (sar > sr) sar = sr;
    double dds = fmax(L.sq + R.sq, 1e-14);
    double vyds = (L.sq * L.vys + R.sq * R.vys + (R.bys - L.bys) * sbx) / dds;
    double vzds = (L.sq * L.vzs + R.sq * R.vzs + (R.bzs - L.bzs) * sbx) / dds;
    double byds = (L.sq * R.bys + R.sq * L.bys + L.sq * R.sq * (R.vys - L.vys) * sbx) / dds;
    double bzds = (L.sq * R.bzs + R.sq * L.bzs + L.sq * R.sq * (R.vzs - L.vzs) * sbx) / dds;
    double vbds = sm * bxf + vyds * byds + vzds * bzds;
    L.eds = L.es - L.sq * (L.vbs - vbds) * sbx;
    R.eds = R.es + R.sq * (R.vbs - vbds) * sbx;

    /* flux selection, physics.py:177-269 */
    if (sl >= 0.0) {
        side_flux(&L, bxf, bx2, psif, ch, f);
    } else if (sr <= 0.0) {
        side_flux(&R, bxf, bx2, psif, ch, f);
    } else {
        double fl[NV], fr[NV], ul[NV], uls[NV], ulds[NV], ur[NV], urs[NV], urds[NV];
        double fls[NV], frs[NV], flds[NV], frds[NV];
        side_flux(&L, bxf, bx2, psif, ch, fl);
        side_flux(&R, bxf, bx2, psif, ch, fr);
        side_states(&L, bxf, psif, sm, vyds, vzds, byds, bzds, ul, uls, ulds);
        side_states(&R, bxf, psif, sm, vyds, vzds, byds, bzds, ur, urs, urds);
        for (int v = 0; v < NV; ++v) {
            fls[v] = fl[v] + sl * (uls[v] - ul[v]);
            frs[v] = fr[v] + sr * (urs[v] - ur[v]);
        }
        for (int v = 0; v < NV; ++v) {
            flds[v] = fls[v] + sal * (ulds[v] - uls[v]);
            frds[v] = frs[v] + sar * (urds[v] - urs[v]);
        }
        const double* pick = (sal >= 0.0) ? fls : (sm >= 0.0) ? flds : (sar >= 0.0) ? frds : frs;
        for (int v = 0; v < NV; ++v) f[v] = pick[v];
        if (ch > 0.0) {
            f[4] = psif;
            f[8] = ch * ch * bxf;
        }
    }
    return smax;
}

/* physics.py:353-379.  WL, WR, F: (NV, n) planes.  dir 0: x faces;
 * dir 1: y faces via the (vx,vy),(Bx,By) swap of physics.py:373-379.
 * smax may be NULL; otherwise receives max over faces of max(|SL|,|SR|). */
void orc_hlld_flux(const double* WL, const double* WR, double* F, double* smax_out,
                   long n, int dir, double gamma, double ch, double floor_rho, double floor_pr)
{
    static const int perm_x[NV] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    static const int perm_y[NV] = {0, 2, 1, 3, 5, 4, 6, 7, 8, 9};
    const int* perm = dir ? perm_y : perm_x;
    double smax = 0.0;
    #pragma omp parallel for schedule(static) reduction(max:smax)
    for (long k = 0; k < n; ++k) {
        double wl[NV], wr[NV], f[NV];
        for (int v = 0; v < NV; ++v) { wl[v] = WL[perm[v] * n + k]; wr[v] = WR[perm[v] * n + k]; }
        double s = hlld_face(wl, wr, f, gamma, ch, floor_rho, floor_pr);
        if (s > smax) smax = s;
        for (int v = 0; v < NV; ++v) F[perm[v] * n + k] = f[v];
    }
    if (smax_out) *smax_out = smax;
}

/* ---------------------------------------------------------------- */
/* solver.py:158-197, in place on U (NV, nxt, nyt). */
void orc_apply_bc(double* U, int nx, int ny, int bcx_periodic, int bcy_periodic)
{
    int nxt = nx + 2 * NG, nyt = ny + 2 * NG;
    long plane = (long)nxt * nyt;
    for (int v = 0; v < NV; ++v) {
        double* u = U + v * plane;
        /* x first, over the full y range */
        for (int g = 0; g < NG; ++g) {
            int lo_dst = g, hi_dst = nxt - NG + g;
            int lo_src, hi_src;
            if (bcx_periodic) { lo_src = nxt - 2 * NG + g; hi_src = NG + g; }
            else { lo_src = NG; hi_src = nxt - NG - 1; }
            memcpy(u + (long)lo_dst * nyt, u + (long)lo_src * nyt, sizeof(double) * nyt);
            memcpy(u + (long)hi_dst * nyt, u + (long)hi_src * nyt, sizeof(double) * nyt);
        }
        if (!bcx_periodic && v == Q_PSI) {
            for (int g = 0; g < NG; ++g) {
                memset(u + (long)g * nyt, 0, sizeof(double) * nyt);
                memset(u + (long)(nxt - NG + g) * nyt, 0, sizeof(double) * nyt);
            }
        }
        /* then y, over the full x range */
        for (int i = 0; i < nxt; ++i) {
            double* r = u + (long)i * nyt;
            for (int g = 0; g < NG; ++g) {
                if (bcy_periodic) { r[g] = r[nyt - 2 * NG + g]; r[nyt - NG + g] = r[NG + g]; }
                else { r[g] = r[NG]; r[nyt - NG + g] = r[nyt - NG - 1]; }
            }
        }
    }
}

/* solver.py:199-204, on n cells per plane. */
void orc_enforce_bounds(double* U, long n)
{
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        double rho = dmax(U[Q_RHO * n + k], FLOOR_RHO);
        double c = U[Q_RHOC * n + k] / rho;
        c = c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
        U[Q_RHOC * n + k] = rho * c;
    }
}

/* solver.py:206-223, over all n cells of W.  Returns dt; *ch_out = glm_ch. */
double orc_compute_dt(const double* W, long n, double gamma, double cfl, double dx, double dy,
                      double* ch_out)
{
    double cfmax = 0.0, vmax = 0.0;
    #pragma omp parallel for schedule(static) reduction(max:cfmax, vmax)
    for (long k = 0; k < n; ++k) {
        double rho = dmax(W[P_RHO * n + k], FLOOR_RHO);
        double p = dmax(W[P_PR * n + k], FLOOR_PR);
        double bx = W[P_BX * n + k], by = W[P_BY * n + k], bz = W[P_BZ * n + k];
        double b2 = bx * bx + by * by + bz * bz;
        double cf = sqrt(gamma * p / rho + b2 / rho);
        double va = fabs(W[P_VX * n + k]) + fabs(W[P_VY * n + k]);
        if (cf > cfmax) cfmax = cf;
        if (va > vmax) vmax = va;
    }
    double ch = dmax(cfmax + vmax, 1.0) * 1.5;
    *ch_out = ch;
    double sm = 0.0;
    #pragma omp parallel for schedule(static) reduction(max:sm)
    for (long k = 0; k < n; ++k) {
        double rho = dmax(W[P_RHO * n + k], FLOOR_RHO);
        double p = dmax(W[P_PR * n + k], FLOOR_PR);
        double bx = W[P_BX * n + k], by = W[P_BY * n + k], bz = W[P_BZ * n + k];
        double b2 = bx * bx + by * by + bz * bz;
        double cf = sqrt(gamma * p / rho + b2 / rho);
        double sx = fabs(W[P_VX * n + k]) + dmax(cf, ch);
        double sy = fabs(W[P_VY * n + k]) + dmax(cf, ch);
        double inv = dmax(sx / dx, sy / dy);
        if (inv > sm) sm = inv;
    }
    if (sm < 1e-14) return cfl * (dx < dy ? dx : dy);
    return cfl / sm;
}

/* solver.py:225-240.  U ghosted; sums in j order. */
double orc_boundary_flux(const double* U, const orc_params* p)
{
    if (p->bcx_periodic) return 0.0;
    int nxt = p->nx + 2 * NG, nyt = p->ny + 2 * NG;
    long plane = (long)nxt * nyt;
    double out[2];
    int cols[2] = {NG, nxt - NG - 1};
    for (int s = 0; s < 2; ++s) {
        double acc = 0.0;
        for (int j = NG; j < nyt - NG; ++j) {
            long k = (long)cols[s] * nyt + j;
            double u[NV], w[NV];
            for (int v = 0; v < NV; ++v) u[v] = U[v * plane + k];
            c2p_cell(u, w, p->gamma);
            double b2 = w[P_BX] * w[P_BX] + w[P_BY] * w[P_BY] + w[P_BZ] * w[P_BZ];
            double pt = w[P_PR] + 0.5 * b2;
            double vb = w[P_VX] * w[P_BX] + w[P_VY] * w[P_BY] + w[P_VZ] * w[P_BZ];
            acc += ((u[Q_EN] + pt) * w[P_VX] - w[P_BX] * vb) * p->dy;
        }
        out[s] = acc;
    }
    return out[0] - out[1];
}

/* solver.py:242-266.  W ghosted -> S ghosted (zero outside [1:-1,1:-1]
 * for divB; S itself is formed everywhere from divB). */
void orc_powell_source(const double* W, double* S, const orc_params* p)
{
    int nxt = p->nx + 2 * NG, nyt = p->ny + 2 * NG;
    long plane = (long)nxt * nyt;
    const double* bx = W + P_BX * plane;
    const double* by = W + P_BY * plane;
    memset(S, 0, sizeof(double) * NV * plane);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nxt; ++i) {
        for (int j = 0; j < nyt; ++j) {
            long k = (long)i * nyt + j;
            double divb = 0.0;
            if (i >= 1 && i < nxt - 1 && j >= 1 && j < nyt - 1) {
                divb = (0.5 * (bx[k] + bx[k + nyt]) - 0.5 * (bx[k - nyt] + bx[k])) / p->dx
                     + (0.5 * (by[k] + by[k + 1]) - 0.5 * (by[k - 1] + by[k])) / p->dy;
            }
            double vx = W[P_VX * plane + k], vy = W[P_VY * plane + k], vz = W[P_VZ * plane + k];
            double Bx = W[P_BX * plane + k], By = W[P_BY * plane + k], Bz = W[P_BZ * plane + k];
            double vb = vx * Bx + vy * By + vz * Bz;
            S[Q_MX * plane + k] = -divb * Bx;
            S[Q_MY * plane + k] = -divb * By;
            S[Q_MZ * plane + k] = -divb * Bz;
            S[Q_BX * plane + k] = -divb * vx;
            S[Q_BY * plane + k] = -divb * vy;
            S[Q_BZ * plane + k] = -divb * vz;
            S[Q_EN * plane + k] = -divb * vb;
        }
    }
}

/* solver.py:268-308.  Applies BC to U in place (solver.py:276); R ghosted,
 * zero in ghosts. */
void orc_compute_rhs(double* U, double* R, const orc_params* p, double ch)
{
    int nx = p->nx, ny = p->ny, nxt = nx + 2 * NG, nyt = ny + 2 * NG;
    long plane = (long)nxt * nyt;
    orc_apply_bc(U, nx, ny, p->bcx_periodic, p->bcy_periodic);
    double* W = (double*)malloc(sizeof(double) * NV * plane);
    orc_cons_to_prim(U, W, plane, p->gamma);

    long nfx = (long)(nxt - 3) * nyt, nfy = (long)nxt * (nyt - 3);
    double* wl = (double*)malloc(sizeof(double) * NV * (nfx > nfy ? nfx : nfy));
    double* wr = (double*)malloc(sizeof(double) * NV * (nfx > nfy ? nfx : nfy));
    double* fx = (double*)malloc(sizeof(double) * NV * nfx);
    double* fy = (double*)malloc(sizeof(double) * NV * nfy);
    orc_muscl_x(W, nxt, nyt, wl, wr);
    orc_hlld_flux(wl, wr, fx, NULL, nfx, 0, p->gamma, ch, FLOOR_RHO, FLOOR_PR);
    orc_muscl_y(W, nxt, nyt, wl, wr);
    orc_hlld_flux(wl, wr, fy, NULL, nfy, 1, p->gamma, ch, FLOOR_RHO, FLOOR_PR);

    double* S = NULL;
    if (p->powell) {
        S = (double*)malloc(sizeof(double) * NV * plane);
        orc_powell_source(W, S, p);
    }
    memset(R, 0, sizeof(double) * NV * plane);
    int nyf = nyt - 3;
    #pragma omp parallel for schedule(static)
    for (int i = NG; i < NG + nx; ++i) {
        for (int v = 0; v < NV; ++v) {
            for (int j = NG; j < NG + ny; ++j) {
                /* Fx index: face k = i-1 (left face of cell i is face i-2... ) --
                   dFx[:, c, :] = Fx[:, c+1, :] - Fx[:, c, :] for interior column
                   c = i - NG; x face c lies left of ghosted cell c + NG. */
                int c = i - NG, r = j - NG;
                double dfx = fx[v * nfx + (long)(c + 1) * nyt + j] - fx[v * nfx + (long)c * nyt + j];
                double dfy = fy[v * nfy + (long)i * nyf + (r + 1)] - fy[v * nfy + (long)i * nyf + r];
                long k = (long)i * nyt + j;
                double acc = 0.0;
                acc -= dfx / p->dx;
                acc -= dfy / p->dy;
                if (S) acc += S[v * plane + k];
                R[v * plane + k] = acc;
            }
        }
    }
    free(W); free(wl); free(wr); free(fx); free(fy); free(S);
}

/* solver.py:310-333.  U in/out (ghosted).  Returns the boundary energy
 * flux of solver.py:316 (0 for periodic x).  decay: exp(-alpha*ch*dt/min(dx,dy))
 * computed by the caller (the reference evaluates it on the host). */
double orc_step_ssprk3(double* U, const orc_params* p, double dt, double ch, double decay)
{
    int nxt = p->nx + 2 * NG, nyt = p->ny + 2 * NG;
    long plane = (long)nxt * nyt, n = NV * plane;
    double* u0 = (double*)malloc(sizeof(double) * n);
    double* u1 = (double*)malloc(sizeof(double) * n);
    double* r = (double*)malloc(sizeof(double) * n);
    memcpy(u0, U, sizeof(double) * n);
    double bflux = orc_boundary_flux(u0, p);

    orc_compute_rhs(u0, r, p, ch);
    for (long k = 0; k < n; ++k) u1[k] = u0[k] + dt * r[k];
    orc_enforce_bounds(u1, plane);
    orc_apply_bc(u1, p->nx, p->ny, p->bcx_periodic, p->bcy_periodic);

    orc_compute_rhs(u1, r, p, ch);
    for (long k = 0; k < n; ++k) U[k] = 0.75 * u0[k] + 0.25 * (u1[k] + dt * r[k]);
    orc_enforce_bounds(U, plane);
    orc_apply_bc(U, p->nx, p->ny, p->bcx_periodic, p->bcy_periodic);

    memcpy(u1, U, sizeof(double) * n);      /* u1 now holds U2 */
    orc_compute_rhs(u1, r, p, ch);
    for (long k = 0; k < n; ++k) U[k] = (1.0 / 3.0) * u0[k] + (2.0 / 3.0) * (u1[k] + dt * r[k]);
    orc_enforce_bounds(U, plane);
    orc_apply_bc(U, p->nx, p->ny, p->bcx_periodic, p->bcy_periodic);

    if (ch > 0) {
        for (long k = 0; k < plane; ++k) U[Q_PSI * plane + k] *= decay;
    }
    free(u0); free(u1); free(r);
    return bflux;
}

/* Time loop of solver.py:475-482 (diagnostics excluded).  Runs until
 * t >= t_end or steps >= max_steps (or dt <= 1e-16).  The psi decay uses
 * libm exp here; the product evaluates it the same way on the host side
 * (numpy) for the step-by-step API. */
int orc_run(double* U, const orc_params* p, double cfl, double t_end, int max_steps,
            double* t_inout, double* bflux_cum, double* last_dt)
{
    int nxt = p->nx + 2 * NG, nyt = p->ny + 2 * NG;
    long plane = (long)nxt * nyt;
    double* W = (double*)malloc(sizeof(double) * NV * plane);
    double t = *t_inout;
    int steps = 0;
    double mind = p->dx < p->dy ? p->dx : p->dy;
    while (t < t_end && steps < max_steps) {
        double ch;
        orc_cons_to_prim(U, W, plane, p->gamma);
        double dt = orc_compute_dt(W, plane, p->gamma, cfl, p->dx, p->dy, &ch);
        if (t_end - t < dt) dt = t_end - t;
        if (dt <= 1e-16) break;
        double decay = exp(-p->glm_alpha * ch * dt / mind);
        double bf = orc_step_ssprk3(U, p, dt, ch, decay);
        if (bflux_cum) *bflux_cum += bf * dt;
        if (last_dt) *last_dt = dt;
        t += dt;
        steps += 1;
    }
    free(W);
    *t_inout = t;
    return steps;
}
