// amhd_device.hpp -- device-side building blocks of the MHD update on gfx950.
//
// Everything here is templated on the arithmetic type `real` and a math
// policy `P`:
//   * ExactMath<double>: the reference's IEEE arithmetic (true division,
//     correctly rounded sqrt, reference epsilons).  Compiled with
//     -ffp-contract=off this reproduces the reference's FP64 per-operation
//     results bit for bit.
//   * FastMath<float>: FP32 with hardware reciprocal/sqrt (v_rcp_f32,
//     v_sqrt_f32, ~1 ulp), hoisted common reciprocals and epsilons scaled to
//     FP32 resolution.  Its error is measured against the FP64 oracle.
//
// Variable order (config.py:13-22):
//   conserved [rho, mx, my, mz, Bx, By, Bz, E, psi, rhoC]
//   primitive [rho, vx, vy, vz, Bx, By, Bz, p, psi, C]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amhd {

constexpr int NV = 10;
enum : int { RHO = 0, MX, MY, MZ, BX, BY, BZ, EN, PSI, RHOC };
enum : int { iRHO = 0, iVX, iVY, iVZ, iBX, iBY, iBZ, iPR, iPSI, iCLR };
constexpr double FLOOR_RHO = 1e-12;  // config.py:40
constexpr double FLOOR_PR = 1e-12;   // config.py:41

// ------------------------------------------------------------------ policies
template <class real> struct ExactMath {
    static constexpr bool exact = true;
    static __device__ __forceinline__ real rcp(real b) { return real(1) / b; }
    // a / b; `ib` (a precomputed reciprocal) is ignored so the quotient is the
    // correctly rounded IEEE one.
    static __device__ __forceinline__ real divi(real a, real b, real /*ib*/) { return a / b; }
    static __device__ __forceinline__ real div(real a, real b) { return a / b; }
    static __device__ __forceinline__ real sqrt_(real a) { return sqrt(a); }
    // Reciprocals of a pair; only the fast policy uses them.
    static __device__ __forceinline__ void rcp2(real a, real b, real& ia, real& ib) { ia = real(1) / a; ib = real(1) / b; }
    // sqrt(x) and 1/sqrt(x) for x >= floor (> 0); exact: only sqrt is used.
    static __device__ __forceinline__ void sqrt_rsq(real x, real& sq, real& rsq) { sq = sqrt(x); rsq = real(0); }
    // epsilons of physics.py:114,117-118,122-123,130
    static __device__ __forceinline__ real eps_den() { return real(1e-14); }
    static __device__ __forceinline__ real eps_sm(real /*sl*/, real /*sr*/) { return real(1e-14); }
    static __device__ __forceinline__ real eps_rel_scale() { return real(1e-8); }
};

template <class real> struct FastMath;
template <> struct FastMath<float> {
    static constexpr bool exact = false;
    static __device__ __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float divi(float a, float /*b*/, float ib) { return a * ib; }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
    // The reciprocals of a pair.  Two v_rcp_f32 measured faster on gfx950
    // (about +0.4% on the whole stage kernel) than one reciprocal of the
    // product and three multiplies (1/a = b / (a b), 1/b = a / (a b)), and
    // the product cannot overflow or underflow.
    static __device__ __forceinline__ void rcp2(float a, float b, float& ia, float& ib) {
        ia = __builtin_amdgcn_rcpf(a);
        ib = __builtin_amdgcn_rcpf(b);
    }
    static __device__ __forceinline__ void sqrt_rsq(float x, float& sq, float& rsq) {
        rsq = __builtin_amdgcn_rsqf(x);
        sq = x * rsq;
    }
    // FP32 cannot resolve the reference's absolute 1e-14 next to O(1..10)
    // speeds: SL + 1e-14 == SL.  The SM clamp therefore uses a margin of
    // 2^-20 of the fan width, and the degenerate-denominator guard a
    // relative 1e-5 instead of 1e-8.
    static __device__ __forceinline__ float eps_den() { return 1e-14f; }
    static __device__ __forceinline__ float eps_sm(float sl, float sr) {
        return fmaxf(1e-14f, 9.5367431640625e-07f * (sr - sl));
    }
    static __device__ __forceinline__ float eps_rel_scale() { return 1e-5f; }
};

// ------------------------------------------------------------- wave shifts
// Whole-wavefront DPP shifts (gfx9 wave_shr:1 / wave_shl:1).  Lane l of
// from_prev() receives lane l-1's value; lane l of from_next() receives lane
// l+1's.  Lane 0 / lane 63 receive 0 (they are halo lanes and never output).
__device__ __forceinline__ float from_prev(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_next(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ double from_prev(double x) {
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x138, 0xf, 0xf, true);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_next(double x) {
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x130, 0xf, 0xf, true);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Max over the 64 lanes of a wave (result valid in every lane).
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}
// Integer max over the 64 lanes, returned as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_max_int(int v) {
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ double wave_sum(double v) {
    // Fixed butterfly order: deterministic for a given lane assignment.
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------- index remap
// Ghost handling of solver.py:158-197 as an index map onto interior cells:
// periodic -> wrap, otherwise zero-gradient copy of the nearest interior
// cell.  Valid for any integer i (the y halo lanes may run far past ny).
__device__ __forceinline__ int remap(int i, int n, int periodic) {
    if (periodic) {
        int r = i % n;
        return r < 0 ? r + n : r;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// --------------------------------------------------------------- primitives
// physics.py:284-301 for one cell.  `psi_zero`: x-ghost of a non-periodic x
// boundary, where solver.py:174-176 sets psi = 0.
template <class real, class P>
__device__ __forceinline__ void cons_to_prim(const real (&u)[NV], real (&w)[NV], real gamma, bool psi_zero) {
    real rho;
    if constexpr (P::exact) rho = u[RHO] > real(FLOOR_RHO) ? u[RHO] : real(FLOOR_RHO);
    else rho = fmax(u[RHO], real(FLOOR_RHO));
    const real ri = P::rcp(rho);
    real vx = u[MX] * ri, vy = u[MY] * ri, vz = u[MZ] * ri;
    real bx = u[BX], by = u[BY], bz = u[BZ];
    w[iRHO] = rho; w[iVX] = vx; w[iVY] = vy; w[iVZ] = vz;
    w[iBX] = bx; w[iBY] = by; w[iBZ] = bz;
    w[iPSI] = psi_zero ? real(0) : u[PSI];
    w[iCLR] = u[RHOC] * ri;
    if constexpr (P::exact) {
        real ke = real(0.5) * rho * (vx * vx + vy * vy + vz * vz);
        real me = real(0.5) * (bx * bx + by * by + bz * bz);
        real p = (gamma - real(1)) * (u[EN] - ke - me);
        w[iPR] = p > real(FLOOR_PR) ? p : real(FLOOR_PR);
    } else {
        // rho v^2 = m.v (v = m / rho); p = (gamma - 1) (E - (m.v + B^2) / 2)
        const real s = u[MX] * vx + u[MY] * vy + u[MZ] * vz + bx * bx + by * by + bz * bz;
        w[iPR] = fmax((gamma - real(1)) * (u[EN] - real(0.5) * s), real(FLOOR_PR));
    }
}

// van Leer limited slope, physics.py:325-328, from the one-sided
// differences dm = w0 - wm, dp = wp - w0.
template <class real, class P>
__device__ __forceinline__ real vl_slope_d(real dm, real dp) {
    const real ab = dm * dp;
    const real d = P::div(real(2) * ab, dm + dp + real(1e-30));
    return ab > real(0) ? d : real(0);
}

template <class real, class P>
__device__ __forceinline__ real vl_slope(real wm, real w0, real wp) {
    return vl_slope_d<real, P>(w0 - wm, wp - w0);
}

// Half the van Leer slope: the offset of the face states W +/- slope/2
// (physics.py:329-336) from the cell value.  The fast path forms
// ab / (dm + dp) directly, which equals 0.5 * (2 ab / (..)) exactly (scaling
// by 2 and by 1/2 is exact), and drops the 1e-30 guard: where ab > 0, dm
// and dp are nonzero with one sign, so dm + dp != 0.  The select ab > 0 ?
// .. : 0 becomes max(ab, 0) * (1 / (dm + dp)) with the legacy multiply,
// for which 0 * x = 0 for every x (inf and NaN included): a flat stencil
// (dm = dp = 0, reciprocal inf) or an extremum gives 0, and ab > 0 gives
// the ordinary IEEE product.
// (v_mul_legacy_f32 through its LLVM intrinsic, so that the compiler
// schedules it and inserts the wait states after a v_rcp_f32 it consumes)
extern "C" __device__ float amhd_fmul_legacy(float, float) __asm("llvm.amdgcn.fmul.legacy");
__device__ __forceinline__ float mul_legacy(float a, float b) { return amhd_fmul_legacy(a, b); }

template <class real, class P>
__device__ __forceinline__ real vl_half_d(real dm, real dp) {
    if constexpr (P::exact) {
        return real(0.5) * vl_slope_d<real, P>(dm, dp);
    } else {
        const real ab = dm * dp;
        return mul_legacy(fmaxf(ab, 0.0f), P::rcp(dm + dp));
    }
}

// ---------------------------------------------------------------- HLLD face
// Constants of one flux evaluation; derived once per launch so that every
// per-face quotient the reference evaluates (0.5/ch, ch*ch) has the same
// IEEE value.
template <class real> struct FluxConst {
    real gamma, gm1, igm1;  // gamma, gamma-1, 1/(gamma-1) (fast path only)
    real ch, half_over_ch, half_ch, ch2;
    real floor_rho, floor_pr;
};

template <class real>
__host__ __device__ inline FluxConst<real> make_flux_const(double gamma, double ch, double floor_rho = FLOOR_RHO,
                                                           double floor_pr = FLOOR_PR) {
    FluxConst<real> c;
    c.floor_rho = real(floor_rho);
    c.floor_pr = real(floor_pr);
    c.gamma = real(gamma);
    c.gm1 = real(gamma - 1.0);
    c.igm1 = real(1.0 / (gamma - 1.0));
    c.ch = real(ch);
    c.half_over_ch = ch > 0.0 ? real(0.5 / ch) : real(0);
    c.half_ch = real(0.5 * ch);
    c.ch2 = real(ch * ch);
    return c;
}

// The HLLD x-flux of physics.py:43-275 for one face.  wl, wr are primitive
// vectors (for y faces the caller passes them with vx<->vy, Bx<->By swapped
// and swaps f back, physics.py:373-379).  The reference evaluates all four
// intermediate fluxes and picks one; here the side K of the contact that
// holds the face (K = L iff SM >= 0, since the clamps keep SAL <= SM <= SAR)
// is selected first and only F_K, F*_K and F**_K are formed.  Each value is
// computed with the reference's operation order, so the FP64 result is
// identical.  Returns max(|SL|, |SR|).
//
// GLM = true: the caller guarantees ch > 0 (always true in the time loop,
// where ch = 1.5 max(cf + |v|, 1) >= 1.5).  Then SL <= -ch < 0 < ch <= SR,
// every face lies in the star fan, and f[4], f[8] are the GLM values
// (physics.py:265-268); the supersonic branches are compiled out.
//
// FLOORED = true: the caller has already floored rho and p of both states
// with the same comparison (physics.py:329-336 floors the MUSCL face
// states), so the floors of physics.py:63-72 would return their inputs
// unchanged (NaN included) and are skipped.
template <class real, class P, bool GLM = false, bool FLOORED = false>
__device__ __forceinline__ real hlld_face(const real (&wl)[NV], const real (&wr)[NV], real (&f)[NV],
                                          const FluxConst<real>& k) {
    const real half = real(0.5);
    real rhoL = (!FLOORED && wl[0] < k.floor_rho) ? k.floor_rho : wl[0];
    real rhoR = (!FLOORED && wr[0] < k.floor_rho) ? k.floor_rho : wr[0];
    real vxL = wl[1], vyL = wl[2], vzL = wl[3];
    real vxR = wr[1], vyR = wr[2], vzR = wr[3];
    real bxL = wl[4], byL = wl[5], bzL = wl[6];
    real bxR = wr[4], byR = wr[5], bzR = wr[6];
    real pL = (!FLOORED && wl[7] < k.floor_pr) ? k.floor_pr : wl[7];
    real pR = (!FLOORED && wr[7] < k.floor_pr) ? k.floor_pr : wr[7];
    real psiL = wl[8], psiR = wr[8];
    real cL = wl[9], cR = wr[9];
    const bool glm = GLM || k.ch > real(0);

    // GLM face Bx / psi (physics.py:74-82)
    real bxf, psif;
    if (glm) {
        bxf = half * (bxL + bxR) - k.half_over_ch * (psiR - psiL);
        psif = half * (psiL + psiR) - k.half_ch * (bxR - bxL);
    } else {
        bxf = half * (bxL + bxR);
        psif = real(0);
    }
    const real bx2 = bxf * bxf;
    const real bmL2 = bx2 + byL * byL + bzL * bzL;
    const real bmR2 = bx2 + byR * byR + bzR * bzR;
    const real ptL = pL + half * bmL2;
    const real ptR = pR + half * bmR2;

    // fast magnetosonic speeds (physics.py:90-102)
    real cfL, cfR;
    if constexpr (P::exact) {
        real irL, irR;
        P::rcp2(rhoL, rhoR, irL, irR);
        real aL2 = P::divi(k.gamma * pL, rhoL, irL);
        real aR2 = P::divi(k.gamma * pR, rhoR, irR);
        real tmpL = aL2 + P::divi(bmL2, rhoL, irL);
        real tmpR = aR2 + P::divi(bmR2, rhoR, irR);
        real discL = tmpL * tmpL - P::divi(real(4) * aL2 * bx2, rhoL, irL);
        real discR = tmpR * tmpR - P::divi(real(4) * aR2 * bx2, rhoR, irR);
        if (discL < real(0)) discL = real(0);
        if (discR < real(0)) discR = real(0);
        cfL = P::sqrt_(fmax(half * (tmpL + P::sqrt_(discL)), real(0)));
        cfR = P::sqrt_(fmax(half * (tmpR + P::sqrt_(discR)), real(0)));
    } else {
        // the same speed with the 1/rho factored out: with A = gamma p,
        // B = |B|^2, X = Bx^2, cf^2 = (A + B + sqrt((A + B)^2 - 4 A X)) / (2 rho);
        // A + B + sqrt(..) >= 0, so the outer clamp is not needed
        // 1 / (2 rho) of both sides from one reciprocal: r = 1 / (2 rhoL rhoR)
        const real hr = half * P::rcp(rhoL * rhoR);
        const real hirL = rhoR * hr, hirR = rhoL * hr;
        const real aL = k.gamma * pL, aR = k.gamma * pR;
        const real sL = aL + bmL2, sR = aR + bmR2;
        const real dL = fmax(sL * sL - real(4) * aL * bx2, real(0));
        const real dR = fmax(sR * sR - real(4) * aR * bx2, real(0));
        cfL = P::sqrt_((sL + P::sqrt_(dL)) * hirL);
        cfR = P::sqrt_((sR + P::sqrt_(dR)) * hirR);
    }

    real sl = fmin(vxL - cfL, vxR - cfR);
    real sr = fmax(vxL + cfL, vxR + cfR);
    if (glm) { sl = fmin(sl, -k.ch); sr = fmax(sr, k.ch); }
    const real smax = fmax(fabs(sl), fabs(sr));

    // contact speed (physics.py:111-120)
    const real dsL = sl - vxL, dsR = sr - vxR;
    real den = dsR * rhoR - dsL * rhoL;
    if constexpr (P::exact) {
        if (fabs(den) < P::eps_den()) den = (den >= real(0)) ? P::eps_den() : -P::eps_den();
    } else {
        // dsR >= cfR >= 0 >= dsL (SR >= vxR + cfR and SL <= vxL - cfL as
        // rounded), so den >= 0 and the clamp above is max(den, eps)
        den = fmax(den, P::eps_den());
    }
    real sm = P::div(dsR * rhoR * vxR - dsL * rhoL * vxL - ptR + ptL, den);
    {
        const real e = P::eps_sm(sl, sr);
        if constexpr (P::exact) {
            if (sm < sl + e) sm = sl + e;
            if (sm > sr - e) sm = sr - e;
        } else {
            // one median: SL + e <= SR - e, as e = max(1e-14, 2^-20 (SR - SL)) and
            // SR - SL >= cf_L + cf_R > 0
            sm = __builtin_amdgcn_fmed3f(sm, sl + e, sr - e);
        }
    }
    const real pts = ptL + rhoL * dsL * (sm - vxL);

    // star states (physics.py:122-154)
    real dslsm = sl - sm, dsrsm = sr - sm;
    if constexpr (P::exact || !GLM) {
        if (fabs(dslsm) < P::eps_den()) dslsm = P::eps_den();
        if (fabs(dsrsm) < P::eps_den()) dsrsm = P::eps_den();
    }
    // (fast path with GLM: SL <= -ch < 0 < ch <= SR gives SR - SL >= |SL|, |SR|,
    // so the SM margin 2^-20 (SR - SL) exceeds an ulp of SL and SR: SM differs
    // from both by at least one ulp, |SL - SM|, |SR - SM| >= 2^-24 ch >> 1e-14,
    // and the clamps cannot fire)
    real idslsm, idsrsm;
    P::rcp2(dslsm, dsrsm, idslsm, idsrsm);
    const real rhoLS = fmax(P::divi(rhoL * dsL, dslsm, idslsm), k.floor_rho);
    const real rhoRS = fmax(P::divi(rhoR * dsR, dsrsm, idsrsm), k.floor_rho);
    real sqLS, sqRS, rsqLS, rsqRS;
    P::sqrt_rsq(rhoLS, sqLS, rsqLS);
    P::sqrt_rsq(rhoRS, sqRS, rsqRS);

    const real eps_rel = P::eps_rel_scale() * fmax(fabs(pts), real(1));
    real dnL = rhoL * dsL * dslsm - bx2;
    real dnR = rhoR * dsR * dsrsm - bx2;
    if constexpr (P::exact) {
        if (fabs(dnL) < eps_rel) dnL = (dnL >= real(0)) ? eps_rel : -eps_rel;
        if (fabs(dnR) < eps_rel) dnR = (dnR >= real(0)) ? eps_rel : -eps_rel;
    } else {
        // the same guard as max and a sign copy (it differs only for D = -0)
        dnL = copysignf(fmaxf(fabsf(dnL), eps_rel), dnL);
        dnR = copysignf(fmaxf(fabsf(dnR), eps_rel), dnR);
    }
    real idnL, idnR;
    P::rcp2(dnL, dnR, idnL, idnR);

    real vyLS, vzLS, byLS, bzLS, vyRS, vzRS, byRS, bzRS;
    if constexpr (P::exact) {
        vyLS = vyL - P::divi(bxf * byL * (sm - vxL), dnL, idnL);
        vzLS = vzL - P::divi(bxf * bzL * (sm - vxL), dnL, idnL);
        byLS = P::divi(byL * (rhoL * dsL * dsL - bx2), dnL, idnL);
        bzLS = P::divi(bzL * (rhoL * dsL * dsL - bx2), dnL, idnL);
        vyRS = vyR - P::divi(bxf * byR * (sm - vxR), dnR, idnR);
        vzRS = vzR - P::divi(bxf * bzR * (sm - vxR), dnR, idnR);
        byRS = P::divi(byR * (rhoR * dsR * dsR - bx2), dnR, idnR);
        bzRS = P::divi(bzR * (rhoR * dsR * dsR - bx2), dnR, idnR);
    } else {
        // the factors shared by the two transverse components, once per side
        const real tL = bxf * (sm - vxL) * idnL, tR = bxf * (sm - vxR) * idnR;
        const real gL = (rhoL * dsL * dsL - bx2) * idnL, gR = (rhoR * dsR * dsR - bx2) * idnR;
        vyLS = vyL - byL * tL; vzLS = vzL - bzL * tL;
        byLS = byL * gL; bzLS = bzL * gL;
        vyRS = vyR - byR * tR; vzRS = vzR - bzR * tR;
        byRS = byR * gR; bzRS = bzR * gR;
    }

    // region: supersonic L / R, or the star fan split by the contact.  Only
    // the side K that holds the face needs its total energy, star energy,
    // v.B and Alfven speed; the double-star state needs both star states.
    const bool supL = !GLM && sl >= real(0);
    const bool supR = !GLM && !supL && sr <= real(0);
    const bool star = GLM || (!supL && !supR);
    const bool useL = supL || (star && sm >= real(0));

    const real rho = useL ? rhoL : rhoR;
    const real vx = useL ? vxL : vxR;
    const real vy = useL ? vyL : vyR;
    const real vz = useL ? vzL : vzR;
    const real by = useL ? byL : byR;
    const real bz = useL ? bzL : bzR;
    const real pk = useL ? pL : pR;
    const real bmk2 = useL ? bmL2 : bmR2;
    const real pt = useL ? ptL : ptR;
    const real psi = useL ? psiL : psiR;
    const real c = useL ? cL : cR;
    const real sk = useL ? sl : sr;
    const real dksm = useL ? dslsm : dsrsm;
    const real idksm = useL ? idslsm : idsrsm;
    const real rhos = useL ? rhoLS : rhoRS;
    const real sqk = useL ? sqLS : sqRS;
    const real rsqk = useL ? rsqLS : rsqRS;
    const real vys = useL ? vyLS : vyRS;
    const real vzs = useL ? vzLS : vzRS;
    const real bys = useL ? byLS : byRS;
    const real bzs = useL ? bzLS : bzRS;

    // (physics.py:141-154 for side K)
    const real vb = vx * bxf + vy * by + vz * bz;
    const real vbs = sm * bxf + vys * bys + vzs * bzs;
    const real e = P::divi(pk, k.gm1, k.igm1) + half * rho * (vx * vx + vy * vy + vz * vz) + half * bmk2;
    const real es = P::divi((sk - vx) * e - pt * vx + pts * sm + bxf * (vb - vbs), dksm, idksm);

    // double-star state (physics.py:156-172)
    const real sbx = (bxf >= real(0)) ? real(1) : real(-1);
    const real abx = fabs(bxf);
    // S*_K = SM -/+ |Bx| / max(sqrt(rho*_K), 1e-14), clamped into [SL, SM] / [SM, SR];
    // rho* >= floor_rho so the fast path multiplies by 1/sqrt(rho*)
    const real tk = P::divi(abx, fmax(sqk, real(1e-14)), rsqk);
    real sak;
    if constexpr (P::exact) {
        if (useL) {
            sak = sm - tk;
            if (sak < sl) sak = sl;
            if (sak > sm) sak = sm;
        } else {
            sak = sm + tk;
            if (sak < sm) sak = sm;
            if (sak > sr) sak = sr;
        }
    } else {
        // the same clamps as one median: SL <= SM - t <= SM <= SR for K = L and
        // SL <= SM <= SM + t for K = R (t >= 0), so min(max(SM - t, SL), SM) and
        // min(max(SM + t, SM), SR) are both med3(SM -/+ t, SL, SR)
        sak = __builtin_amdgcn_fmed3f(sm + (useL ? -tk : tk), sl, sr);
    }
    // physical flux of side K (physics.py:178-187) and U_K, U*_K
    // (physics.py:213-249)
    auto side_flux = [&](real (&fk)[NV], real (&uk)[NV], real (&us)[NV]) {
        fk[0] = rho * vx;
        fk[1] = rho * vx * vx + pt - bx2;
        fk[2] = rho * vy * vx - bxf * by;
        fk[3] = rho * vz * vx - bxf * bz;
        fk[4] = glm ? psif : real(0);
        fk[5] = by * vx - bxf * vy;
        fk[6] = bz * vx - bxf * vz;
        fk[7] = (e + pt) * vx - bxf * vb;
        fk[8] = glm ? bxf * k.ch * k.ch : real(0);
        fk[9] = rho * c * vx;
        uk[0] = rho; uk[1] = rho * vx; uk[2] = rho * vy; uk[3] = rho * vz; uk[4] = bxf;
        uk[5] = by; uk[6] = bz; uk[7] = e; uk[8] = psi; uk[9] = rho * c;
        us[0] = rhos; us[1] = rhos * sm; us[2] = rhos * vys; us[3] = rhos * vzs; us[4] = bxf;
        us[5] = bys; us[6] = bzs; us[7] = es; us[8] = psif; us[9] = rhos * c;
    };

    if constexpr (P::exact) {
        // F = F_K outside the star fan, F*_K = F_K + S_K (U*_K - U_K) in it,
        // and F**_K = F*_K + S*_K (U**_K - U*_K) in the double-star region
        // (physics.py:178-275).  U** is formed only where a face is in that
        // region, with the same operations as the reference on the same
        // values, so every face's flux is the same bit for bit.
        const bool dbl = star && (useL ? (sak < real(0)) : (sak >= real(0)));
        real fk[NV], uk[NV], us[NV];
        side_flux(fk, uk, us);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (GLM && (v == 4 || v == 8)) continue;   // set below
            const real fs = fk[v] + sk * (us[v] - uk[v]);
            f[v] = star ? fs : fk[v];
        }
        if (dbl) {
            // double-star state (physics.py:156-172)
            const real dds = fmax(sqLS + sqRS, real(1e-14));
            const real idds = P::rcp(dds);
            const real vyDS = P::divi(sqLS * vyLS + sqRS * vyRS + (byRS - byLS) * sbx, dds, idds);
            const real vzDS = P::divi(sqLS * vzLS + sqRS * vzRS + (bzRS - bzLS) * sbx, dds, idds);
            const real byDS = P::divi(sqLS * byRS + sqRS * byLS + sqLS * sqRS * (vyRS - vyLS) * sbx, dds, idds);
            const real bzDS = P::divi(sqLS * bzRS + sqRS * bzLS + sqLS * sqRS * (vzRS - vzLS) * sbx, dds, idds);
            const real vbDS = sm * bxf + vyDS * byDS + vzDS * bzDS;
            // E** (physics.py:171-172): L side subtracts, R side adds
            const real eds = useL ? (es - sqk * (vbs - vbDS) * sbx) : (es + sqk * (vbs - vbDS) * sbx);
            const real uds[NV] = {rhos, rhos * sm, rhos * vyDS, rhos * vzDS, bxf, byDS, bzDS, eds, psif, rhos * c};
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if (GLM && (v == 4 || v == 8)) continue;   // set below
                f[v] = f[v] + sak * (uds[v] - us[v]);   // dbl implies star: f[v] is F*_K
            }
        }
    } else {
        // F* = F_K + S_K (U*_K - U_K) (S_K = 0 outside the star fan), then
        // F** = F* + S*_K (U**_K - U*_K) for the faces in the double-star
        // region only; components 0, 1, 4, 8, 9 of U** equal those of U*.
        // A wavefront with no double-star face skips the second part.
        real fk[NV], uk[NV], us[NV];
        side_flux(fk, uk, us);
        const real sks = star ? sk : real(0);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (GLM && (v == 4 || v == 8)) continue;   // set below
            f[v] = fk[v] + sks * (us[v] - uk[v]);
        }
        if (star) {
            // the mass and scalar components of F_K + S_K (U*_K - U_K) in
            // fewer operations: rho*_K (S_K - SM) = rho_K (S_K - v_K) (the
            // star density, physics.py:122-125) gives rho v + S (rho* - rho) =
            // rho* SM, and C is the same in U_K and U*_K; the same flux up to
            // rounding (and the density floor, which only a vacuum meets)
            const real ms = rhos * sm;
            f[0] = ms;
            f[9] = ms * c;
        }
        // S*_K < 0 (K = L) or >= 0 (K = R).  At S*_K = 0 the correction
        // below vanishes, so both sides may leave it out there: the test is
        // (K = L ? -S*_K : S*_K) > 0, and a face with Bx = 0 and SM = 0 (a
        // uniform state at rest) skips the branch
        const bool dbl = star && (useL ? -sak : sak) > real(0);
        if (dbl) {
            // double-star state (physics.py:156-172) and E** (physics.py:171-172)
            const real dds = fmax(sqLS + sqRS, real(1e-14));
            const real idds = P::rcp(dds);
            const real vyDS = P::divi(sqLS * vyLS + sqRS * vyRS + (byRS - byLS) * sbx, dds, idds);
            const real vzDS = P::divi(sqLS * vzLS + sqRS * vzRS + (bzRS - bzLS) * sbx, dds, idds);
            const real byDS = P::divi(sqLS * byRS + sqRS * byLS + sqLS * sqRS * (vyRS - vyLS) * sbx, dds, idds);
            const real bzDS = P::divi(sqLS * bzRS + sqRS * bzLS + sqLS * sqRS * (vzRS - vzLS) * sbx, dds, idds);
            const real vbDS = sm * bxf + vyDS * byDS + vzDS * bzDS;
            const real eds = useL ? (es - sqk * (vbs - vbDS) * sbx) : (es + sqk * (vbs - vbDS) * sbx);
            const real uds[NV] = {rhos, rhos * sm, rhos * vyDS, rhos * vzDS, bxf, byDS, bzDS, eds, psif, rhos * c};
#pragma unroll
            for (int v = 2; v < 8; ++v) {
                if (v == 4) continue;
                f[v] = f[v] + sak * (uds[v] - us[v]);
            }
        }
    }
    if (glm) {  // physics.py:265-268 (every glm face is in the star fan)
        f[4] = psif;
        f[8] = k.ch2 * bxf;
    }
    return smax;
}

// The ideal-MHD x-flux of one state with the GLM components (psi, ch^2 Bx):
// the HLLD flux of a face whose two states are equal (its consistency
// F(W, W) = F(W); with WL = WR the GLM face values are Bx_f = Bx and
// psi_f = psi, physics.py:74-82).  FP32 fast path only.
template <class real>
__device__ __forceinline__ void phys_flux_x(const real (&w)[NV], real (&f)[NV], const FluxConst<real>& k) {
    const real rho = w[iRHO], vx = w[iVX], vy = w[iVY], vz = w[iVZ];
    const real bx = w[iBX], by = w[iBY], bz = w[iBZ], p = w[iPR];
    const real half = real(0.5);
    const real b2 = bx * bx + by * by + bz * bz;
    const real pt = p + half * b2;
    const real m = rho * vx;
    const real e = p * k.igm1 + half * (rho * (vx * vx + vy * vy + vz * vz) + b2);
    f[0] = m;
    f[1] = m * vx + pt - bx * bx;
    f[2] = m * vy - bx * by;
    f[3] = m * vz - bx * bz;
    f[4] = w[iPSI];
    f[5] = by * vx - bx * vy;
    f[6] = bz * vx - bx * vz;
    f[7] = (e + pt) * vx - bx * (vx * bx + vy * by + vz * bz);
    f[8] = k.ch2 * bx;
    f[9] = m * w[iCLR];
}

// True when a and b hold the same bits in every component.  A computation
// fed with the same bits returns the same bits, so a face whose inputs are
// those of a face already solved may take that face's result.
__device__ __forceinline__ bool same_bits(const float (&a)[NV], const float (&b)[NV]) {
    unsigned acc = 0u;
#pragma unroll
    for (int v = 0; v < NV; ++v) acc |= __float_as_uint(a[v]) ^ __float_as_uint(b[v]);
    return acc == 0u;
}
__device__ __forceinline__ bool same_bits(const double (&a)[NV], const double (&b)[NV]) {
    unsigned long long acc = 0ull;
#pragma unroll
    for (int v = 0; v < NV; ++v)
        acc |= (unsigned long long)__double_as_longlong(a[v]) ^ (unsigned long long)__double_as_longlong(b[v]);
    return acc == 0ull;
}

// y-face flux: rotate (vx,vy),(Bx,By), solve, rotate back (physics.py:373-379)
template <class real, class P, bool GLM = false, bool FLOORED = false>
__device__ __forceinline__ real hlld_face_y(const real (&wl)[NV], const real (&wr)[NV], real (&f)[NV],
                                            const FluxConst<real>& k) {
    real a[NV] = {wl[0], wl[2], wl[1], wl[3], wl[5], wl[4], wl[6], wl[7], wl[8], wl[9]};
    real b[NV] = {wr[0], wr[2], wr[1], wr[3], wr[5], wr[4], wr[6], wr[7], wr[8], wr[9]};
    real g[NV];
    real s = hlld_face<real, P, GLM, FLOORED>(a, b, g, k);
    f[0] = g[0]; f[1] = g[2]; f[2] = g[1]; f[3] = g[3]; f[4] = g[5];
    f[5] = g[4]; f[6] = g[6]; f[7] = g[7]; f[8] = g[8]; f[9] = g[9];
    return s;
}

}  // namespace amhd
