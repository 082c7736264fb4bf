// Stockham FFT axis passes with fused split-step operations, for gfx950.
//
// One launch transforms every pencil of one lattice axis.  A workgroup of 256
// threads (four wave64s) owns a tile of W pencils x N points:
//   phase A  global -> LDS, coalesced, through a "load op" (pointwise physics,
//            imprint, rescale, reductions) applied to every element;
//   phase B  Stockham stages in LDS (radix 16/8/4/3/2 butterflies in
//            registers, self-sorting so the result is in natural order);
//   phase C  optional "mid op" on the transformed tile (kinetic phase,
//            k^2 factor, potential half steps, reductions) and optional second
//            transform in the other direction (turn-around passes);
//   phase D  LDS -> global, coalesced, or nothing when only reductions are
//            wanted.
// The z axis is contiguous (pencil tiles are contiguous runs); y and x are
// strided, with W consecutive z-columns per tile so every global row access
// is a W-element contiguous run.
//
// Transform convention (the reference's cuFFT Z2Z, orginal/gpe_engine.cu:519-527):
//   X[k] = sum_j x[j] exp(sign 2 pi i j k / N), sign -1 forward, +1 inverse,
// unnormalised.
#pragma once

#include "gpe_common.hpp"

namespace gpe {

constexpr int kThreads = 256;

// ------------------------------------------------------------ FFT plans

struct FftPlan {
    int ns;
    int r[8];
};

constexpr FftPlan make_plan(int N) {
    FftPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    switch (N) {  // hand-picked schedules for the common lengths
    case 16: p.ns = 1; p.r[0] = 16; return p;
    case 32: p.ns = 2; p.r[0] = 8; p.r[1] = 4; return p;
    case 64: p.ns = 2; p.r[0] = 8; p.r[1] = 8; return p;
    case 128: p.ns = 2; p.r[0] = 16; p.r[1] = 8; return p;
    case 256: p.ns = 2; p.r[0] = 16; p.r[1] = 16; return p;
    case 512: p.ns = 3; p.r[0] = 8; p.r[1] = 8; p.r[2] = 8; return p;
    default: break;
    }
    int n = N;
    while (n % 16 == 0 && n > 16) { p.r[p.ns++] = 16; n /= 16; }
    while (n % 8 == 0) { p.r[p.ns++] = 8; n /= 8; }
    while (n % 4 == 0) { p.r[p.ns++] = 4; n /= 4; }
    while (n % 2 == 0) { p.r[p.ns++] = 2; n /= 2; }
    while (n % 3 == 0) { p.r[p.ns++] = 3; n /= 3; }
    if (n == 16) p.r[p.ns++] = 16;
    return p;
}

constexpr int plan_span(const FftPlan& p, int s) {
    int v = 1;
    for (int i = 0; i < s; i++) v *= p.r[i];
    return v;
}

constexpr bool plan_ok(int N) {
    FftPlan p = make_plan(N);
    return p.ns > 0 && plan_span(p, p.ns) == N;
}

// tile width (pencils per workgroup), a power of two
template <typename T, int N>
constexpr int tile_w() {
    int base = (sizeof(T) == 4) ? 4096 : 2048;
    int w = base / N;
    if (w < 1) w = 1;
    int p2 = 1;
    while (p2 * 2 <= w) p2 *= 2;
    if (p2 > kThreads) p2 = kThreads;
    return p2;
}

template <int N, int W, bool CONTIG>
__device__ __forceinline__ int lidx(int w, int i) {
    if constexpr (CONTIG)
        return w * (N + N / 16) + i + (i >> 4);
    else
        return i * (W + 1) + w;
}
template <int N, int W, bool CONTIG>
constexpr int lds_elems() {
    return CONTIG ? W * (N + N / 16) : N * (W + 1);
}

// ---------------------------------------------------- radix butterflies

// exact cos / sin of 2 pi j / 16
struct Roots16 {
    static constexpr double c[16] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                                     0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
                                     -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977,
                                     0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674};
    static constexpr double s[16] = {0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674,
                                     1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                                     0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
                                     -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977};
};

// a * exp(S 2 pi i E / R) for R dividing 16
template <int R, int E, int S, typename T>
__device__ __forceinline__ cplx<T> rot(cplx<T> a) {
    constexpr int e = E % R;
    constexpr int j = e * (16 / R);
    if constexpr (e == 0) {
        return a;
    } else if constexpr (j == 4) {  // exp(S i pi/2) = S i
        return (S > 0) ? mk(-a.y, a.x) : mk(a.y, -a.x);
    } else if constexpr (j == 8) {
        return mk(-a.x, -a.y);
    } else if constexpr (j == 12) {  // exp(S i 3pi/2) = -S i
        return (S > 0) ? mk(a.y, -a.x) : mk(-a.y, a.x);
    } else {
        constexpr double cv = Roots16::c[j];
        constexpr double sv = S * Roots16::s[j];
        if constexpr (sizeof(T) == 4) {
            // complex64: the constant as a float hi + lo pair, so that its
            // rounding does not shrink |a| the same way on every application
            constexpr float ch = (float)cv, cl = (float)(cv - (double)ch);
            constexpr float sh = (float)sv, sl = (float)(sv - (double)sh);
            return mk(fmaf(a.x, ch, fmaf(-a.y, sh, a.x * cl - a.y * sl)),
                      fmaf(a.x, sh, fmaf(a.y, ch, a.x * sl + a.y * cl)));
        } else {
            const T c = (T)cv;
            const T s = (T)sv;
            return mk(a.x * c - a.y * s, a.x * s + a.y * c);
        }
    }
}

// a * (hi + lo) for a complex64 twiddle stored as a hi + lo pair
__device__ __forceinline__ cplx<float> cmul_comp(cplx<float> a, cplx<float> hi, cplx<float> lo) {
    return mk(fmaf(a.x, hi.x, fmaf(-a.y, hi.y, a.x * lo.x - a.y * lo.y)),
              fmaf(a.x, hi.y, fmaf(a.y, hi.x, a.x * lo.y + a.y * lo.x)));
}

template <int R, int S, typename T>
__device__ __forceinline__ void dft(cplx<T>* v);

template <int S, typename T>
__device__ __forceinline__ void dft2(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int S, typename T>
__device__ __forceinline__ void dft3(cplx<T>* v) {
    constexpr double hd = 0.86602540378443864676;  // sqrt(3)/2
    cplx<T> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    cplx<T> m = mk(v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y);
    cplx<T> ht2;
    if constexpr (sizeof(T) == 4) {
        constexpr float hh = (float)hd, hl = (float)(hd - (double)hh);
        ht2 = mk(fmaf(t2.x, hh, t2.x * hl), fmaf(t2.y, hh, t2.y * hl));
    } else {
        ht2 = mk((T)hd * t2.x, (T)hd * t2.y);
    }
    cplx<T> it2 = mk(-(T)S * ht2.y, (T)S * ht2.x);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, it2);
    v[2] = csub(m, it2);
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cplx<T>* v) {
    cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cplx<T> t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    cplx<T> it3 = (S > 0) ? mk(-t3.y, t3.x) : mk(t3.y, -t3.x);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, it3);
    v[3] = csub(t1, it3);
}

template <int R1, int R2, int S, typename T, int I>
__device__ __forceinline__ void apply_inner_twiddles(cplx<T> (&y)[R2][R1]) {
    if constexpr (I < R1 * R2) {
        y[I / R1][I % R1] = rot<R1 * R2, (I / R1) * (I % R1), S>(y[I / R1][I % R1]);
        apply_inner_twiddles<R1, R2, S, T, I + 1>(y);
    }
}

// R = R1 * R2 with m = R2*m1 + m2 and k = k1 + R1*k2:
// X[k] = sum_m2 w_R^(m2 k1) [sum_m1 x[R2 m1 + m2] w_R1^(m1 k1)] w_R2^(m2 k2)
template <int R1, int R2, int S, typename T>
__device__ __forceinline__ void dft_composite(cplx<T>* v) {
    constexpr int R = R1 * R2;
    cplx<T> y[R2][R1];
#pragma unroll
    for (int m2 = 0; m2 < R2; m2++) {
#pragma unroll
        for (int m1 = 0; m1 < R1; m1++) y[m2][m1] = v[R2 * m1 + m2];
        dft<R1, S>(y[m2]);
    }
    apply_inner_twiddles<R1, R2, S, T, 0>(y);  // w_R^(m2 k1), compile-time
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
        cplx<T> z[R2];
#pragma unroll
        for (int m2 = 0; m2 < R2; m2++) z[m2] = y[m2][k1];
        dft<R2, S>(z);
#pragma unroll
        for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = z[k2];
    }
}

template <int R, int S, typename T>
__device__ __forceinline__ void dft(cplx<T>* v) {
    if constexpr (R == 2) dft2<S>(v);
    else if constexpr (R == 3) dft3<S>(v);
    else if constexpr (R == 4) dft4<S>(v);
    else if constexpr (R == 8) dft_composite<4, 2, S>(v);
    else if constexpr (R == 16) dft_composite<4, 4, S>(v);
    else static_assert(R == 2, "unsupported radix");
}

// ------------------------------------------------------ Stockham stages

// Stage with radix R after a span NS: butterfly j of pencil w reads
// (w, j + r N/R), applies w_N^(r (j mod NS) N/(NS R)), transforms, and writes
// (w, (j/NS) NS R + j mod NS + r NS).
template <typename T, int N, int W, bool CONTIG, int S, int R, int NS>
__device__ __forceinline__ void fft_stage(cplx<T>* lds, const cplx<T>* tw, const cplx<T>* twlo) {
    constexpr int NB = W * N / R;
    constexpr int PER = (NB + kThreads - 1) / kThreads;
    constexpr int M = N / R;
    constexpr int TSTEP = N / (NS * R);
    const int tid = threadIdx.x;
    cplx<T> v[PER][R];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = tid + q * kThreads;
        if ((NB % kThreads) == 0 || b < NB) {
            int w, j;
            if constexpr (CONTIG) { w = b / M; j = b % M; }
            else { j = b / W; w = b % W; }
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = lds[lidx<N, W, CONTIG>(w, j + r * M)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = tid + q * kThreads;
        if ((NB % kThreads) == 0 || b < NB) {
            int w, j;
            if constexpr (CONTIG) { w = b / M; j = b % M; }
            else { j = b / W; w = b % W; }
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    cplx<T> t = tw[k * r * TSTEP];
                    if (S > 0) t.y = -t.y;
                    if constexpr (sizeof(T) == 4) {
                        cplx<T> tl = twlo[k * r * TSTEP];
                        if (S > 0) tl.y = -tl.y;
                        v[q][r] = cmul_comp(v[q][r], t, tl);
                    } else {
                        v[q][r] = cmul(v[q][r], t);
                    }
                }
            }
            dft<R, S>(v[q]);
            const int idxD = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[lidx<N, W, CONTIG>(w, idxD + r * NS)] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int W, bool CONTIG, int S, int STAGE>
__device__ __forceinline__ void fft_stages(cplx<T>* lds, const cplx<T>* tw, const cplx<T>* twlo) {
    constexpr FftPlan P = make_plan(N);
    if constexpr (STAGE < P.ns) {
        fft_stage<T, N, W, CONTIG, S, P.r[STAGE], plan_span(P, STAGE)>(lds, tw, twlo);
        fft_stages<T, N, W, CONTIG, S, STAGE + 1>(lds, tw, twlo);
    }
}

// ------------------------------------------------------ pass arguments

enum LoadMode { LOAD_PLAIN = 0, LOAD_PSI0 = 1, LOAD_PART2 = 2 };
enum MidMode {
    MID_NONE = 0,
    MID_EXPT = 1,        // kinetic factor Kx Ky Kz/N (orginal/gpe_engine.cu:880-895)
    MID_K2 = 2,          // qf' |k|^2 (orginal/gpe_engine.cu:897-909)
    MID_VSTEP2 = 3,      // predictor end + corrector potential (orginal 774-801 / 817-868)
    MID_PART1 = 4,       // predictor end only, norm sum (orginal 803-815)
    MID_VSTEP3 = 5,      // corrector end (orginal 870-878), optional next Vstep1
    MID_QF_VSTEP1 = 6,   // QF overlap (orginal 911-922) + Vstep1_qf (orginal 748-772)
    MID_KIN_SUM = 7,     // sum |k|^2 |psi_k|^2 (kinetic energy by Parseval)
    MID_KA = 8,          // k_a / (gamma N) (orginal 1125-1162)
    MID_OVERLAP_RE = 9,  // Re(conj psi * v) -> real output (orginal 1067-1074)
};
enum StoreMode { STORE_PLAIN = 0, STORE_NONE = 1 };
enum PassFlags {
    F_SCALE = 1,          // multiply loaded psi0 / phi3 by the device normalisation factor
    F_IMPRINT = 2,        // imprint the vortex on the loaded psi0 (written back)
    F_VSTEP1 = 4,         // LOAD_PSI0: apply Vstep1 after the preparation
    F_QF = 8,             // add the stored quantum-friction potential
    F_RED_POT = 16,       // LOAD_PLAIN: sum rho*Vext and EDF(rho)
    F_RED_NORM = 32,      // sum rho of the produced wave function
    F_NEXT_VSTEP1 = 64,   // MID_VSTEP3: continue with the next step's Vstep1
    F_NEXT_PSI = 128,     // MID_VSTEP3: continue with the next step's psi0 (QF transform)
    F_IMPRINT_NEXT = 256, // MID_VSTEP3: imprint the next step's psi0
    F_IMPRINT_MID = 512,  // LOAD_PART2: re-imprint psi0 before the corrector (ITE vortex)
};

template <typename T>
struct PassArgs {
    int axis;              // 0 x, 1 y, 2 z
    unsigned M;            // number of pencils
    int dir1, dir2;        // -1 / +1; dir2 == 0: single transform
    int load_mode, mid_mode, store_mode, flags;
    int nred;              // number of FP64 sums produced (0..3)
    const cplx<T>* src;
    cplx<T>* dst;
    const cplx<T>* tw;     // exp(-2 pi i m / N), m < N
    const cplx<T>* twlo;   // complex64: exp(-2 pi i m / N) - tw, the low part
    cplx<T>* psiA;         // the wave function buffer (psi0 / psi1)
    T* vbuf;               // corrector potential V1 per point
    T* vq;                 // quantum-friction potential per point
    double* rout;          // real FP64 output (currents)
    const cplx<double>* Kx;  // kinetic factors in FP64, Kz includes 1/N
    const cplx<double>* Ky;
    const cplx<double>* Kz;
    const T* kx;           // wave numbers
    const T* ky;
    const T* kz;
    const double* kx64;    // wave numbers in FP64 for the energy sums
    const double* ky64;
    const double* kz64;
    const double* scal;    // device scalars; scal[1] is the normalisation factor
    double* partials;      // [nred][gridDim.x]
    T kfac;                // MID_K2 / MID_KA factor
    int ka_axis;           // MID_KA axis
    double up64[kMaxUserParams];  // user params in FP64 (energy density)
    Phys<T> ph;
};

struct Coord {
    unsigned off, ix, iy, iz;
};

template <typename T>
__device__ __forceinline__ Coord coord(const PassArgs<T>& a, int N, unsigned p, unsigned i) {
    Coord c;
    const unsigned ny = (unsigned)a.ph.ny, nz = (unsigned)a.ph.nz;
    if (a.axis == 2) {
        c.ix = p / ny;
        c.iy = p - c.ix * ny;
        c.iz = i;
        c.off = p * (unsigned)N + i;
    } else if (a.axis == 1) {
        c.ix = p / nz;
        c.iz = p - c.ix * nz;
        c.iy = i;
        c.off = c.ix * (ny * nz) + c.iz + i * nz;
    } else {
        c.iy = p / nz;
        c.iz = p - c.iy * nz;
        c.ix = i;
        c.off = p + i * (ny * nz);
    }
    return c;
}

// v * exp(re + i im).  In complex64 the factor's angle and modulus come from
// single-precision sincos/exp, the direction is renormalised to exact unit
// modulus and the product taken in FP64, so the rounding of the factor does
// not bias |psi| the same way every step.
__device__ __forceinline__ cplx<double> mul_exp(cplx<double> v, double re, double im) {
    return cmul(v, cexp_(re, im));
}
__device__ __forceinline__ cplx<float> mul_exp(cplx<float> v, float re, float im) {
    float s, c;
    sincosf(im, &s, &c);
    const double cd = (double)c, sd = (double)s;
    const double f = (double)expf(re) / sqrt(cd * cd + sd * sd);
    const double vx = (double)v.x, vy = (double)v.y;
    return mk((float)(f * (vx * cd - vy * sd)), (float)(f * (vx * sd + vy * cd)));
}

template <typename T>
__device__ __forceinline__ T pot_at(const PassArgs<T>& a, const Coord& c, cplx<T> v, bool qf) {
    T lv = vext(a.ph, c.ix, c.iy, c.iz, c.off) + dedfdn(a.ph, density(a.ph, v));
    if (qf) lv = lv + a.vq[c.off];
    return lv;
}

// Vstep1: psi -> modify(psi) exp(c V(psi))  (orginal/gpe_engine.cu:723-772)
template <typename T>
__device__ __forceinline__ cplx<T> vstep1(const PassArgs<T>& a, const Coord& c, cplx<T> v, bool qf, T qval) {
    T lv = vext(a.ph, c.ix, c.iy, c.iz, c.off) + dedfdn(a.ph, density(a.ph, v));
    if (qf) lv = lv + qval;
    return mul_exp(v, a.ph.c.x * lv, a.ph.c.y * lv);
}

__device__ __forceinline__ double dens64(double g, double x, double y) { return g * (x * x + y * y); }

template <typename T>
__device__ __forceinline__ cplx<T> load_op(const PassArgs<T>& a, const Coord& c, double* acc) {
    const bool qf = (a.flags & F_QF) != 0;
    if (a.load_mode == LOAD_PLAIN) {
        cplx<T> v = a.src[c.off];
        if (a.flags & F_RED_POT) {
            const double rho = dens64((double)a.ph.gamma, (double)v.x, (double)v.y);
            acc[0] += rho * (double)vext(a.ph, c.ix, c.iy, c.iz, c.off);
            acc[1] += edf(a.ph.interactions, a.up64, rho);
        }
        return v;
    }
    if (a.load_mode == LOAD_PSI0) {
        cplx<T> v = a.psiA[c.off];
        bool wb = false;
        if (a.flags & F_SCALE) { v = cscale(v, (T)a.scal[1]); wb = true; }
        if (a.flags & F_IMPRINT) { v = imprint(a.ph, v, c.ix, c.iy); wb = true; }
        if (wb) a.psiA[c.off] = v;
        if (a.flags & F_VSTEP1) v = vstep1(a, c, v, qf, qf ? a.vq[c.off] : (T)0);
        return v;
    }
    // LOAD_PART2: corrector potential from the (normalised) predictor
    // (orginal/gpe_engine.cu:817-868)
    cplx<T> phi3 = a.src[c.off];
    if (a.flags & F_SCALE) phi3 = cscale(phi3, (T)a.scal[1]);
    cplx<T> psi0 = a.psiA[c.off];
    const T V0 = pot_at(a, c, psi0, qf);
    T lv;
    if (qf)
        lv = (T)0.5 * (V0 + vext(a.ph, c.ix, c.iy, c.iz, c.off) + dedfdn(a.ph, density(a.ph, phi3)) + a.vq[c.off]);
    else
        lv = (T)0.5 * (V0 + vext(a.ph, c.ix, c.iy, c.iz, c.off) + dedfdn(a.ph, density(a.ph, phi3)));
    a.vbuf[c.off] = lv;
    if (a.flags & F_IMPRINT_MID) psi0 = imprint(a.ph, psi0, c.ix, c.iy);
    return mul_exp(psi0, a.ph.c.x * lv, a.ph.c.y * lv);
}

template <typename T>
__device__ __forceinline__ cplx<T> mid_op(const PassArgs<T>& a, const Coord& c, cplx<T> v, double* acc) {
    const bool qf = (a.flags & F_QF) != 0;
    switch (a.mid_mode) {
    case MID_EXPT: {
        // orginal/gpe_engine.cu:880-895, in FP64 in both precisions: the factors
        // are the same every step, so a rounded |K| != 1 would compound
        cplx<double> w = mk((double)v.x, (double)v.y);
        w = cmul(w, a.Kx[c.ix]);
        w = cmul(w, a.Ky[c.iy]);
        w = cmul(w, a.Kz[c.iz]);
        return mk((T)w.x, (T)w.y);
    }
    case MID_K2: {
        const T k2 = a.kx[c.ix] * a.kx[c.ix] + a.ky[c.iy] * a.ky[c.iy] + a.kz[c.iz] * a.kz[c.iz];
        return cscale(v, a.kfac * k2);
    }
    case MID_KA: {
        const T k = (a.ka_axis == 0) ? a.kx[c.ix] : (a.ka_axis == 1) ? a.ky[c.iy] : a.kz[c.iz];
        return cscale(v, k / a.kfac);
    }
    case MID_VSTEP2: {
        // orginal/gpe_engine.cu:774-801 (and the _part1/_part2[_qf] pair for qf)
        const cplx<T> psi0 = a.psiA[c.off];
        const T V0 = pot_at(a, c, psi0, qf);
        const cplx<T> phi3 = mul_exp(v, a.ph.c.x * V0, a.ph.c.y * V0);
        T lv;
        if (qf)
            lv = (T)0.5 * (V0 + vext(a.ph, c.ix, c.iy, c.iz, c.off) + dedfdn(a.ph, density(a.ph, phi3)) + a.vq[c.off]);
        else
            lv = (T)0.5 * (V0 + vext(a.ph, c.ix, c.iy, c.iz, c.off) + dedfdn(a.ph, density(a.ph, phi3)));
        a.vbuf[c.off] = lv;
        return mul_exp(psi0, a.ph.c.x * lv, a.ph.c.y * lv);
    }
    case MID_PART1: {
        // orginal/gpe_engine.cu:803-815 with the density sum of gpe_normalize fused
        const cplx<T> psi0 = a.psiA[c.off];
        const T V0 = pot_at(a, c, psi0, qf);
        const cplx<T> phi3 = mul_exp(v, a.ph.c.x * V0, a.ph.c.y * V0);
        acc[0] += dens64((double)a.ph.gamma, (double)phi3.x, (double)phi3.y);
        return phi3;
    }
    case MID_VSTEP3: {
        // orginal/gpe_engine.cu:870-878; the factor exp(c V1) is recomputed from V1
        const T lv = a.vbuf[c.off];
        cplx<T> psi1 = mul_exp(v, a.ph.c.x * lv, a.ph.c.y * lv);
        if (a.flags & F_RED_NORM) acc[0] += dens64((double)a.ph.gamma, (double)psi1.x, (double)psi1.y);
        if (a.flags & (F_NEXT_VSTEP1 | F_NEXT_PSI)) {
            if (a.flags & F_IMPRINT_NEXT) psi1 = imprint(a.ph, psi1, c.ix, c.iy);
            a.psiA[c.off] = psi1;
            if (a.flags & F_NEXT_VSTEP1) return vstep1(a, c, psi1, false, (T)0);
            return psi1;
        }
        a.psiA[c.off] = psi1;
        return psi1;
    }
    case MID_QF_VSTEP1: {
        // V_qf = Im(conj psi0 * F^-1[qf' k^2 F psi0]) / (rho0 + eps), then Vstep1_qf
        const cplx<T> psi0 = a.psiA[c.off];
        const T r0 = density(a.ph, psi0);
        const T q = (psi0.x * v.y + (-psi0.y) * v.x) / (r0 + (T)1.0e-12);
        a.vq[c.off] = q;
        return vstep1(a, c, psi0, true, q);
    }
    case MID_KIN_SUM: {
        const double k2 = a.kx64[c.ix] * a.kx64[c.ix] + a.ky64[c.iy] * a.ky64[c.iy] + a.kz64[c.iz] * a.kz64[c.iz];
        acc[0] += k2 * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        return v;
    }
    case MID_OVERLAP_RE: {
        const cplx<T> p = a.psiA[c.off];
        a.rout[c.off] = (double)(p.x * v.x + p.y * v.y);
        return v;
    }
    default:
        return v;
    }
}

// FP64 sum over the workgroup, wave64 shuffles then LDS; lane 0 of wave 0 holds the result
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kThreads / 64; i++) s += red[i];
    return s;
}

template <typename T, int N, bool CONTIG>
__global__ __launch_bounds__(kThreads) void fft_pass_kernel(const PassArgs<T> a) {
    constexpr int W = tile_w<T, N>();
    constexpr int TILE = W * N;
    constexpr int PERE = (TILE + kThreads - 1) / kThreads;
    __shared__ cplx<T> lds[lds_elems<N, W, CONTIG>()];
    __shared__ cplx<T> tw[N];
    __shared__ cplx<T> twlo[sizeof(T) == 4 ? N : 1];
    __shared__ double red[kThreads / 64];

    for (int i = threadIdx.x; i < N; i += kThreads) {
        tw[i] = a.tw[i];
        if constexpr (sizeof(T) == 4) twlo[i] = a.twlo[i];
    }
    const unsigned p0 = blockIdx.x * (unsigned)W;
    double acc[3] = {0.0, 0.0, 0.0};

for (int q = 0; q < PERE; q++) {
        const int e = threadIdx.x + q * kThreads;
        if ((TILE % kThreads) == 0 || e < TILE) {
            int w, i;
            if constexpr (CONTIG) { w = e / N; i = e % N; }
            else { i = e / W; w = e % W; }
            const unsigned p = p0 + (unsigned)w;
            cplx<T> v = mk((T)0, (T)0);
            if (p < a.M) v = load_op(a, coord(a, N, p, (unsigned)i), acc);
            lds[lidx<N, W, CONTIG>(w, i)] = v;
        }
    }
    __syncthreads();
    if (a.dir1 < 0) fft_stages<T, N, W, CONTIG, -1, 0>(lds, tw, twlo);
    else fft_stages<T, N, W, CONTIG, 1, 0>(lds, tw, twlo);

    if (a.mid_mode != MID_NONE) {
for (int q = 0; q < PERE; q++) {
            const int e = threadIdx.x + q * kThreads;
            if ((TILE % kThreads) == 0 || e < TILE) {
                int w, i;
                if constexpr (CONTIG) { w = e / N; i = e % N; }
                else { i = e / W; w = e % W; }
                const unsigned p = p0 + (unsigned)w;
                if (p < a.M) {
                    const int li = lidx<N, W, CONTIG>(w, i);
                    lds[li] = mid_op(a, coord(a, N, p, (unsigned)i), lds[li], acc);
                }
            }
        }
        __syncthreads();
    }
    if (a.dir2 < 0) fft_stages<T, N, W, CONTIG, -1, 0>(lds, tw, twlo);
    else if (a.dir2 > 0) fft_stages<T, N, W, CONTIG, 1, 0>(lds, tw, twlo);

    if (a.store_mode == STORE_PLAIN) {
for (int q = 0; q < PERE; q++) {
            const int e = threadIdx.x + q * kThreads;
            if ((TILE % kThreads) == 0 || e < TILE) {
                int w, i;
                if constexpr (CONTIG) { w = e / N; i = e % N; }
                else { i = e / W; w = e % W; }
                const unsigned p = p0 + (unsigned)w;
                if (p < a.M) a.dst[coord(a, N, p, (unsigned)i).off] = lds[lidx<N, W, CONTIG>(w, i)];
            }
        }
    }
    for (int r = 0; r < a.nred; r++) {
        const double s = block_sum(acc[r], red);
        if (threadIdx.x == 0) a.partials[(size_t)r * gridDim.x + blockIdx.x] = s;
    }
}

}  // namespace gpe
