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

// c64: complex64 (512 = 32 x 16: its lean tiles are 16 columns wide, so every
// thread owns a radix-32 or radix-16 butterfly; complex128 keeps 8 x 8 x 8
// for its 4-column tiles)
constexpr FftPlan make_plan(int N, bool c64) {
    FftPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    switch (N) {  // hand-picked schedules for the common lengths
    case 16: p.ns = 1; p.r[0] = 16; return p;
    case 32: p.ns = 2; p.r[0] = 8; p.r[1] = 4; return p;
    case 64: p.ns = 2; p.r[0] = 8; p.r[1] = 8; return p;
    case 128: p.ns = 2; p.r[0] = 16; p.r[1] = 8; return p;
    case 256: p.ns = 2; p.r[0] = 16; p.r[1] = 16; return p;
    case 512:
        if (c64) { p.ns = 2; p.r[0] = 32; p.r[1] = 16; }
        else { p.ns = 3; p.r[0] = 8; p.r[1] = 8; p.r[2] = 8; }
        return p;
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

constexpr bool plan_ok(int N, bool c64) {
    FftPlan p = make_plan(N, c64);
    return p.ns > 0 && plan_span(p, p.ns) == N;
}

// tile width (pencils per workgroup), a power of two
template <typename T, int N>
constexpr int tile_w() {
    // complex64 512: 16 columns (128-byte rows) rather than 8; complex128 128
    // and 256: 32 and 16 columns, so that every thread owns a butterfly of
    // the radix-16 first stage (256-byte rows); complex128 512 keeps 4
    // columns (8 measured slower: 62 against 70 steps/s at 512^3, two
    // workgroups per CU instead of three)
    int base = (sizeof(T) == 4) ? (N >= 512 ? 8192 : 4096) : ((N == 128 || N == 256) ? 4096 : 2048);
    int w = base / N;
    if (w < 1) w = 1;
    int p2 = 1;
    while (p2 * 2 <= w) p2 *= 2;
    if (p2 > kThreads) p2 = kThreads;
    return p2;
}

// Row pitch of a strided tile: rows of 128 bytes or more are stored
// unpadded (a row already spans the lanes that share an LDS cycle, and at
// 256^3 complex64 the saved 2 KB lets four x turn-around workgroups share a
// CU); shorter rows keep one pad element against first-stage write conflicts.
template <typename T, int W>
constexpr int row_pitch() {
    return W * (int)sizeof(cplx<T>) >= 128 ? W : W + 1;
}
template <typename T, int N, int W, bool CONTIG>
__device__ __forceinline__ int lidx(int w, int i) {
    if constexpr (CONTIG)
        return w * (N + N / 16) + i + (i >> 4);
    else
        return i * row_pitch<T, W>() + w;
}
template <typename T, int N, int W, bool CONTIG>
constexpr int lds_elems() {
    return CONTIG ? W * (N + N / 16) : N * row_pitch<T, W>();
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

// exact cos / sin of 2 pi j / 12
struct Roots12 {
    static constexpr double c[12] = {1.0,  0.86602540378443864676,  0.5, 0.0, -0.5, -0.86602540378443864676,
                                     -1.0, -0.86602540378443864676, -0.5, 0.0, 0.5,  0.86602540378443864676};
    static constexpr double s[12] = {0.0,  0.5, 0.86602540378443864676,  1.0,  0.86602540378443864676,  0.5,
                                     0.0,  -0.5, -0.86602540378443864676, -1.0, -0.86602540378443864676, -0.5};
};

// exact cos / sin of 2 pi j / 32 for odd j (entry (j - 1) / 2)
struct Roots32Odd {
    static constexpr double c[16] = {0.98078528040323044913,  0.83146961230254523708,  0.55557023301960222474,
                                     0.19509032201612826785,  -0.19509032201612826785, -0.55557023301960222474,
                                     -0.83146961230254523708, -0.98078528040323044913, -0.98078528040323044913,
                                     -0.83146961230254523708, -0.55557023301960222474, -0.19509032201612826785,
                                     0.19509032201612826785,  0.55557023301960222474,  0.83146961230254523708,
                                     0.98078528040323044913};
    static constexpr double s[16] = {0.19509032201612826785,  0.55557023301960222474,  0.83146961230254523708,
                                     0.98078528040323044913,  0.98078528040323044913,  0.83146961230254523708,
                                     0.55557023301960222474,  0.19509032201612826785,  -0.19509032201612826785,
                                     -0.55557023301960222474, -0.83146961230254523708, -0.98078528040323044913,
                                     -0.98078528040323044913, -0.83146961230254523708, -0.55557023301960222474,
                                     -0.19509032201612826785};
};

// exact cos and sin of 2 pi / 24 (15 degrees): (sqrt 6 + sqrt 2) / 4 and (sqrt 6 - sqrt 2) / 4
constexpr double kCos15 = 0.96592582628906828675, kSin15 = 0.25881904510252076235;

// a * exp(S 2 pi i E / R) for R dividing 32, 24 or 12.  complex64 takes a
// constant as a float hi + lo pair, so that its rounding does not shrink |a|
// the same way on every application
template <int R, int E, int S, typename T>
__device__ __forceinline__ cplx<T> rot(cplx<T> a) {
    if constexpr (R == 3 || R == 6 || R == 12) {
        constexpr int j = (E % R) * (12 / R);
        if constexpr (j == 0) {
            return a;
        } else if constexpr (j == 3) {
            return (S > 0) ? mk(-a.y, a.x) : mk(a.y, -a.x);
        } else if constexpr (j == 6) {
            return mk(-a.x, -a.y);
        } else if constexpr (j == 9) {
            return (S > 0) ? mk(a.y, -a.x) : mk(-a.y, a.x);
        } else {
            constexpr double cv = Roots12::c[j];
            constexpr double sv = S * Roots12::s[j];
            if constexpr (sizeof(T) == 4) {
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
    } else if constexpr (R == 24) {
        constexpr int e = E % 24;
        if constexpr (e % 2 == 0) {
            return rot<12, e / 2, S>(a);
        } else if constexpr (e % 3 == 0) {
            return rot<8, e / 3, S>(a);
        } else {
            // odd multiples of 15 degrees that are not multiples of 45
            constexpr double cv = (e == 1 || e == 23) ? kCos15 : (e == 5 || e == 19) ? kSin15
                                  : (e == 7 || e == 17) ? -kSin15 : -kCos15;
            constexpr double s0 = (e == 1 || e == 11) ? kSin15 : (e == 5 || e == 7) ? kCos15
                                  : (e == 13 || e == 23) ? -kSin15 : -kCos15;
            constexpr double sv = S * s0;
            if constexpr (sizeof(T) == 4) {
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
    } else if constexpr (R == 32) {
        constexpr int e = E % 32;
        if constexpr (e % 2 == 0) {
            return rot<16, e / 2, S>(a);
        } else {
            constexpr double cv = Roots32Odd::c[e / 2];
            constexpr double sv = S * Roots32Odd::s[e / 2];
            if constexpr (sizeof(T) == 4) {
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
    } else {
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
}

// a * (hi + lo) for a complex64 twiddle stored as a hi + lo pair
__device__ __forceinline__ cplx<float> cmul_comp(cplx<float> a, cplx<float> hi, cplx<float> lo) {
    return mk(fmaf(a.x, hi.x, fmaf(-a.y, hi.y, a.x * lo.x - a.y * lo.y)),
              fmaf(a.x, hi.y, fmaf(a.y, hi.x, a.x * lo.y + a.y * lo.x)));
}

// A twiddle entry in LDS: complex64 keeps the float hi and lo parts of
// exp(-2 pi i m / N) together (one 16-byte read), complex128 the value.
template <typename T>
struct TwE;
template <>
struct TwE<float> {
    float4 t;  // hi.x, hi.y, lo.x, lo.y
};
template <>
struct TwE<double> {
    cplx<double> t;
};
// a * w for sign S < 0, a * conj(w) for S > 0
template <int S>
__device__ __forceinline__ cplx<float> tw_mul(cplx<float> a, const TwE<float>& e) {
    const float4 q = e.t;
    return cmul_comp(a, mk(q.x, S > 0 ? -q.y : q.y), mk(q.z, S > 0 ? -q.w : q.w));
}
template <int S>
__device__ __forceinline__ cplx<double> tw_mul(cplx<double> a, const TwE<double>& e) {
    cplx<double> t = e.t;
    if (S > 0) t.y = -t.y;
    return cmul(a, t);
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
    else if constexpr (R == 6) dft_composite<3, 2, S>(v);
    else if constexpr (R == 8) dft_composite<4, 2, S>(v);
    else if constexpr (R == 12) dft_composite<4, 3, S>(v);
    else if constexpr (R == 16) dft_composite<4, 4, S>(v);
    else if constexpr (R == 24) dft_composite<8, 3, S>(v);
    else if constexpr (R == 32) dft_composite<8, 4, S>(v);
    else static_assert(R == 2, "unsupported radix");
}

// ------------------------------------------------------ Stockham stages

// Stage with radix R after a span NS: butterfly j of pencil w reads
// (w, j + r N/R), applies w_N^(r (j mod NS) N/(NS R)), transforms, and writes
// (w, (j/NS) NS R + j mod NS + r NS).
template <typename T, int N, int W, bool CONTIG, int S, int R, int NS>
__device__ __forceinline__ void fft_stage(cplx<T>* lds, const TwE<T>* tw) {
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
            for (int r = 0; r < R; r++) v[q][r] = lds[lidx<T, N, W, CONTIG>(w, j + r * M)];
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
                for (int r = 1; r < R; r++) v[q][r] = tw_mul<S>(v[q][r], tw[k * r * TSTEP]);
            }
            dft<R, S>(v[q]);
            const int idxD = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[lidx<T, N, W, CONTIG>(w, idxD + r * NS)] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int W, bool CONTIG, int S, int STAGE>
__device__ __forceinline__ void fft_stages(cplx<T>* lds, const TwE<T>* tw) {
    constexpr FftPlan P = make_plan(N, sizeof(T) == 4);
    if constexpr (STAGE < P.ns) {
        fft_stage<T, N, W, CONTIG, S, P.r[STAGE], plan_span(P, STAGE)>(lds, tw);
        fft_stages<T, N, W, CONTIG, S, STAGE + 1>(lds, tw);
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
    MID_EXPTC = 10,      // conj(Kx Ky Kz)/N: the adjoint of the kinetic factor
    MID_VPP = 11,        // Vpp(k)/N: dipolar convolution kernel (dev/gpe_engine.cu:960-975)
    MID_VPP_SUM = 12,    // sum Vpp(k) |rho_k|^2 (dev/gpe_engine.cu:1716-1735)
    // the kinetic factor split by axis, K = (K_y K_z) K_x: K_y K_z in a y
    // turn-around whose data is already in k_z, K_x in the x turn-around
    MID_EXPT_SEP = 13,
};

// Fourier transform of the particle-particle potential of z-polarised dipoles
// with contact part, gpe_Vpp_k (dev/gpe_user_defined.h:422-438):
// a_con = up[3], a_dip = up[4]
__device__ __forceinline__ double vpp_k(const double* up, double kx, double ky, double kz) {
    const double PI = 3.14159265358979323846;
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 < 2.220446049250313e-16) return 4. * PI * up[3];
    return 4. * PI * (up[3] + up[4] * (3. * kz * kz / k2 - 1.));
}
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
    unsigned M;            // number of pencils (all instances)
    unsigned n1;           // points per instance (instances are consecutive lattices)
    int dir1, dir2;        // -1 / +1; dir2 == 0: single transform
    int load_mode, mid_mode, store_mode, flags;
    int nred;              // number of FP64 sums produced (0..3)
    int variant;           // launcher choice: bit 0 z passes in LDS tiles, bit 1 half-width strided tiles,
                           // bit 2 pipelined strided passes (complex64)
    unsigned pipe_wgs;     // persistent workgroups of a pipelined pass (2 per CU)
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
    const cplx<float>* kx_hl;   // complex64: K_x as (hi, lo) float pairs
    const cplx<float>* kyz_hl;  // complex64: K_y K_z per (iy, iz) as (hi, lo) pairs
    const double* kx64;    // wave numbers in FP64 for the energy sums
    const double* ky64;
    const double* kz64;
    const double* scal;    // device scalars; scal[1] is the normalisation factor
    double* partials;      // [nred][gridDim.x]
    int pin;               // phase pinning of psi0: 1 vortex, 2 vortex where |psi| > 1e-15, 3 stored phase
    const double* phase;   // stored phase (pin 3)
    T kfac;                // MID_K2 / MID_KA factor
    int ka_axis;           // MID_KA axis
    double up64[kMaxUserParams];  // user params in FP64 (energy density)
    Phys<T> ph;
};

struct Coord {
    unsigned off, ix, iy, iz;
};


// sin and cos of x.  For |x| <= 1/4 (every real-time potential phase
// c V = -dt V / 2 of a sane step) the Taylor polynomials to x^7 and x^6, whose
// truncation error is below 4e-10; otherwise the library sincosf.
__device__ __forceinline__ void sincos_small(float x, float* s, float* c) {
    if (fabsf(x) <= 0.25f) {
        const float t = x * x;
        *s = fmaf(x * t, fmaf(t, fmaf(t, -1.0f / 5040.0f, 1.0f / 120.0f), -1.0f / 6.0f), x);
        *c = fmaf(t, fmaf(t, fmaf(t, -1.0f / 720.0f, 1.0f / 24.0f), -0.5f), 1.0f);
    } else {
        sincosf(x, s, c);
    }
}

// complex128: sin and cos of |x| <= 1/4 from their Taylor series through
// x^13 and x^12 (the first omitted terms are below 1e-17 relative); the
// chunked physics of the z passes uses it, the per-point ops keep sincos
__device__ __forceinline__ void sincos_taylor(double x, double* s, double* c) {
    const double t = x * x;
    double ps = fma(t, -1.0 / 6227020800.0, 1.0 / 39916800.0);
    ps = fma(t, ps, -1.0 / 362880.0);
    ps = fma(t, ps, 1.0 / 5040.0);
    ps = fma(t, ps, -1.0 / 120.0);
    ps = fma(t, ps, 1.0 / 6.0);
    *s = fma(-x * t, ps, x);
    double pc = fma(t, 1.0 / 479001600.0, -1.0 / 3628800.0);
    pc = fma(t, pc, 1.0 / 40320.0);
    pc = fma(t, pc, -1.0 / 720.0);
    pc = fma(t, pc, 1.0 / 24.0);
    pc = fma(t, pc, -0.5);
    *c = fma(t, pc, 1.0);
}

// v * exp(re + i im).  In complex64 the rounded direction (c, s) misses unit
// modulus by e = c^2 + s^2 - 1 ~ 1e-8, the same way every step; e is formed
// without cancellation by two FMAs and the product is scaled by 1 + h,
// h = -e/2 (= 1/|(c, s)| - 1 to O(e^2)).  The scaling must enter before the
// product's one significant rounding: |h| is below half an ulp, so applied to
// an already rounded product it would be rounded away every time.  Hence
//   rx = fma(vx, c, fma(-vy, s, h vx)),  ry = fma(vy, c, fma(vx, s, h vy)),
// whose inner terms are small (s ~ |cV|) and whose dropped part h (r - v) is
// O(h |cV|).  The factor is then a rotation to ~1e-15 up to unbiased roundings.
__device__ __forceinline__ cplx<double> mul_exp(cplx<double> v, double re, double im) {
    return cmul(v, cexp_(re, im));
}
__device__ __forceinline__ cplx<float> mul_exp(cplx<float> v, float re, float im) {
    float s, c;
    sincos_small(im, &s, &c);
    const float h = -0.5f * fmaf(s, s, fmaf(c, c, -1.0f));
    float rx = fmaf(v.x, c, fmaf(-v.y, s, h * v.x));
    float ry = fmaf(v.y, c, fmaf(v.x, s, h * v.y));
    if (re != 0.0f) {  // real time: re == 0 for every point
        const float g = expf(re);
        rx *= g;
        ry *= g;
    }
    return mk(rx, ry);
}


// Side arrays (psi0, V1, V_qf) are written by one z pass and read back by
// the next one four axis passes later; the state array in between is read
// and written by every pass.  The z-wave kernels access them non-temporally,
// so they do not push the state array out of the caches (measured +4-6%
// steps/s at 256^3 complex64, bit-identical results; the LDS-tile kernels of
// the other lengths showed no change and keep ordinary accesses).
template <typename T>
__device__ __forceinline__ void st_nt(cplx<T>* p, cplx<T> v) {
    using V2 = T __attribute__((ext_vector_type(2)));
    V2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<V2*>(p));
}
template <typename T>
__device__ __forceinline__ cplx<T> ld_nt(const cplx<T>* p) {
    using V2 = T __attribute__((ext_vector_type(2)));
    const V2 w = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return mk(w.x, w.y);
}
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
    return __builtin_nontemporal_load(p);
}

// Phase pinning of the ITE variants: __gpe_imprint_vortexline_zdir_ (dev/gpe_engine.cu:485-511),
// __gpe_imprint2_vortexline_zdir_ (520-548) and __gpe_enforce_phase__ (584-591)
template <typename T>
__device__ __forceinline__ cplx<T> apply_pin(const PassArgs<T>& a, cplx<T> v, const Coord& c) {
    if (a.pin == 1) return imprint(a.ph, v, c.ix, c.iy);
    if (a.pin == 2) return (g_hypot(v.x, v.y) > (T)1e-15) ? imprint(a.ph, v, c.ix, c.iy) : v;
    if (a.pin == 3) {
        const T m = g_sqrt(v.x * v.x + v.y * v.y);
        T sn, cs;
        g_sincos((T)a.phase[c.off], &sn, &cs);
        return mk(m * cs, m * sn);
    }
    return v;
}


__device__ __forceinline__ double dens64(double g, double x, double y) { return g * (x * x + y * y); }

// k-space factors of the lean passes, one per compile-time kind; MK_EXPT_HL is
// the complex64 x-axis form of MID_EXPT / MID_EXPTC with the compensated
// tables, MK_EXPT_64 the FP64 product, MK_SEP the per-axis factor of
// MID_EXPT_SEP (AX: the pass axis, -1 when only known at run time)
enum MidKind { MK_EXPT_64 = 0, MK_EXPT_HL = 1, MK_K2 = 2, MK_VPP = 3, MK_KA = 4, MK_SEP = 5 };

// kx: the pass's factor staged in LDS, or null: complex64 the x pass's K_x
// pairs (hi.x, hi.y, lo.x, lo.y), complex128 K_x (x pass) or K_y (y pass)
template <int MK, bool CJ, typename T, int AX = -1>
__device__ __forceinline__ cplx<T> mid_op_lean_k(const PassArgs<T>& a, const Coord& c, cplx<T> v,
                                                 const float4* kx = nullptr) {
    if constexpr (MK == MK_SEP) {
        const bool xpass = (AX >= 0) ? (AX == 0) : (a.axis == 0);
        if constexpr (sizeof(T) == 4) {
            // one compensated product per pass with a hi + lo float pair of
            // an FP64 value: K_x in the x pass, the per-pencil K_y K_z in the
            // y pass (two roundings per kinetic step, as in the unsplit form)
            if (xpass && kx) {
                const float4 q = kx[c.ix];
                return cmul_comp(v, mk(q.x, q.y), mk(q.z, q.w));
            }
            if (xpass) return cmul_comp(v, a.kx_hl[2 * c.ix], a.kx_hl[2 * c.ix + 1]);
            const unsigned yz = c.iy * (unsigned)a.ph.nz + c.iz;
            return cmul_comp(v, a.kyz_hl[2 * yz], a.kyz_hl[2 * yz + 1]);
        } else {
            cplx<double> w = mk((double)v.x, (double)v.y);
            if (kx) {
                // complex128: K_x (x pass) or K_y (y pass) staged in LDS
                w = cmul(w, reinterpret_cast<const cplx<double>*>(kx)[xpass ? c.ix : c.iy]);
                if (!xpass) w = cmul(w, a.Kz[c.iz]);
            } else if (xpass) {
                w = cmul(w, a.Kx[c.ix]);
            } else {
                w = cmul(w, a.Ky[c.iy]);
                w = cmul(w, a.Kz[c.iz]);
            }
            return mk((T)w.x, (T)w.y);
        }
    } else if constexpr (MK == MK_EXPT_HL && sizeof(T) == 4) {
        // complex64 on the x axis: two compensated products with K_x and
        // the per-pencil K_y K_z (hi + lo float pairs of FP64 values)
        const unsigned yz = c.iy * (unsigned)a.ph.nz + c.iz;
        cplx<float> h1 = a.kx_hl[2 * c.ix], l1 = a.kx_hl[2 * c.ix + 1];
        cplx<float> h2 = a.kyz_hl[2 * yz], l2 = a.kyz_hl[2 * yz + 1];
        if (CJ) { h1.y = -h1.y; l1.y = -l1.y; h2.y = -h2.y; l2.y = -l2.y; }
        v = cmul_comp(v, h1, l1);
        return cmul_comp(v, h2, l2);
    } else if constexpr (MK == MK_EXPT_64 || MK == MK_EXPT_HL) {
        cplx<double> w = mk((double)v.x, (double)v.y);
        cplx<double> kx = a.Kx[c.ix], ky = a.Ky[c.iy], kz = a.Kz[c.iz];
        if (CJ) { kx.y = -kx.y; ky.y = -ky.y; kz.y = -kz.y; }
        w = cmul(w, kx);
        w = cmul(w, ky);
        w = cmul(w, kz);
        return mk((T)w.x, (T)w.y);
    } else if constexpr (MK == MK_K2) {
        const T k2 = a.kx[c.ix] * a.kx[c.ix] + a.ky[c.iy] * a.ky[c.iy] + a.kz[c.iz] * a.kz[c.iz];
        return cscale(v, a.kfac * k2);
    } else if constexpr (MK == MK_VPP) {
        const double f = vpp_k(a.up64, a.kx64[c.ix], a.ky64[c.iy], a.kz64[c.iz]) /
                         ((double)a.ph.nx * (double)a.ph.ny * (double)a.ph.nz);
        return mk((T)((double)v.x * f), (T)((double)v.y * f));
    } else {
        const T k = (a.ka_axis == 0) ? a.kx[c.ix] : (a.ka_axis == 1) ? a.ky[c.iy] : a.kz[c.iz];
        return cscale(v, k / a.kfac);
    }
}

// the kind and conjugation of a lean mid op; false when it has none
template <typename T>
__host__ __device__ __forceinline__ bool mid_kind(const PassArgs<T>& a, int& mk, bool& cj) {
    cj = (a.mid_mode == MID_EXPTC);
    switch (a.mid_mode) {
    case MID_EXPT:
    case MID_EXPTC:
        mk = (sizeof(T) == 4 && a.axis == 0 && a.kx_hl) ? MK_EXPT_HL : MK_EXPT_64;
        return true;
    case MID_K2: mk = MK_K2; return true;
    case MID_VPP: mk = MK_VPP; return true;
    case MID_KA: mk = MK_KA; return true;
    case MID_EXPT_SEP: mk = MK_SEP; return true;
    default: return false;
    }
}

template <typename T>
__device__ __forceinline__ cplx<T> mid_op_lean(const PassArgs<T>& a, const Coord& c, cplx<T> v) {
    int mk;
    bool cj;
    if (!mid_kind(a, mk, cj)) return v;
    switch (mk) {
    case MK_EXPT_HL: return cj ? mid_op_lean_k<MK_EXPT_HL, true>(a, c, v) : mid_op_lean_k<MK_EXPT_HL, false>(a, c, v);
    case MK_EXPT_64: return cj ? mid_op_lean_k<MK_EXPT_64, true>(a, c, v) : mid_op_lean_k<MK_EXPT_64, false>(a, c, v);
    case MK_K2: return mid_op_lean_k<MK_K2, false>(a, c, v);
    case MK_VPP: return mid_op_lean_k<MK_VPP, false>(a, c, v);
    case MK_SEP: return mid_op_lean_k<MK_SEP, false>(a, c, v);
    default: return mid_op_lean_k<MK_KA, false>(a, c, v);
    }
}

// Inputs of one element of a physics pass, all taken from LDS (or registers)
// by the caller: the lattice coordinates, the external potential there, psi0
// and the stored real per-point field (V_qf or V1) when the pass needs them.
template <typename T>
struct ElemIn {
    Coord c;
    unsigned it;   // time index of the step (time_index of the pass)
    T vx;          // V_ext at the point at it
    T vx1;         // V_ext at it + 1 (differs from vx only for a time-dependent user potential)
    cplx<T> aux;   // psi0
    T raux;        // V_qf (F_QF) or V1 (MID_VSTEP3)
};

// V(psi) = Vext + dEDFdn(rho(psi)) [+ V_qf] at time index it  (orginal/gpe_engine.cu:739, 764)
template <bool UH, typename T>
__device__ __forceinline__ T pot2(const PassArgs<T>& a, T vx, cplx<T> v, bool qf, T q, unsigned it, unsigned off) {
    T lv = vx + dedfdn<UH>(a.ph, density(a.ph, v), it, off);
    if (qf) lv = lv + q;
    return lv;
}

// gpe_modify_psi at the element (identity unless the engine runs user hooks)
template <bool UH, typename T>
__device__ __forceinline__ cplx<T> modify_at(const PassArgs<T>& a, const ElemIn<T>& e, cplx<T> v, unsigned it) {
    return modify_psi<UH>(a.ph, v, e.c.ix, e.c.iy, e.c.iz, e.c.off, it);
}

// UH: compiled with the user hook file (passes of engines with INTERACTIONS or VEXT 100)
template <int LM, bool UH = false, typename T>
__device__ __forceinline__ cplx<T> load_op_m(const PassArgs<T>& a, const ElemIn<T>& e, cplx<T> v, double* acc) {
    const bool qf = (a.flags & F_QF) != 0;
    const unsigned off = e.c.off;
    if constexpr (LM == LOAD_PLAIN) {
        if (a.flags & F_RED_POT) {
            const double rho = dens64((double)a.ph.gamma, (double)v.x, (double)v.y);
            acc[0] += rho * (double)e.vx;
            acc[1] += edf<UH>(a.ph, a.up64, rho, e.it, off);
        }
        return v;
    } else if constexpr (LM == LOAD_PSI0) {
        // v is psi0 (orginal/gpe_engine.cu:723-772 after the rescale / imprint)
        bool wb = false;
        if (a.flags & F_SCALE) { v = cscale(v, (T)a.scal[1]); wb = true; }
        if (a.flags & F_IMPRINT) { v = apply_pin(a, v, e.c); wb = true; }
        if (wb) a.psiA[off] = v;
        if (a.flags & F_VSTEP1) {
            v = modify_at<UH>(a, e, v, e.it);
            const T lv = pot2<UH>(a, e.vx, v, qf, e.raux, e.it, off);
            v = mul_exp(v, a.ph.c.x * lv, a.ph.c.y * lv);
        }
        return v;
    } else {
        // LOAD_PART2: v is the predictor phi3 (orginal/gpe_engine.cu:817-868)
        cplx<T> phi3 = v;
        if (a.flags & F_SCALE) phi3 = cscale(phi3, (T)a.scal[1]);
        cplx<T> psi0 = e.aux;
        const T V0 = pot2<UH>(a, e.vx, modify_at<UH>(a, e, psi0, e.it), qf, e.raux, e.it, off);
        T lv;
        const T d1 = dedfdn<UH>(a.ph, density(a.ph, phi3), e.it + 1, off);
        if (qf) lv = (T)0.5 * (V0 + e.vx1 + d1 + e.raux);
        else lv = (T)0.5 * (V0 + e.vx1 + d1);
        a.vbuf[off] = lv;
        if (a.flags & F_IMPRINT_MID) psi0 = apply_pin(a, psi0, e.c);
        return mul_exp(modify_at<UH>(a, e, psi0, e.it), a.ph.c.x * lv, a.ph.c.y * lv);
    }
}

template <bool UH, typename T>
__device__ __forceinline__ cplx<T> load_op2(const PassArgs<T>& a, const ElemIn<T>& e, cplx<T> v, double* acc) {
    if (a.load_mode == LOAD_PLAIN) return load_op_m<LOAD_PLAIN, UH>(a, e, v, acc);
    if (a.load_mode == LOAD_PSI0) return load_op_m<LOAD_PSI0, UH>(a, e, v, acc);
    return load_op_m<LOAD_PART2, UH>(a, e, v, acc);
}

template <int MM, bool UH = false, typename T>
__device__ __forceinline__ cplx<T> mid_op_m(const PassArgs<T>& a, const ElemIn<T>& e, cplx<T> v, double* acc) {
    const bool qf = (a.flags & F_QF) != 0;
    const unsigned off = e.c.off;
    if constexpr (MM == MID_VSTEP2) {
        // orginal/gpe_engine.cu:774-801 (and the _part1/_part2[_qf] pair for qf)
        const cplx<T> psi0 = modify_at<UH>(a, e, e.aux, e.it);
        const T V0 = pot2<UH>(a, e.vx, psi0, qf, e.raux, e.it, off);
        const cplx<T> phi3 = mul_exp(v, a.ph.c.x * V0, a.ph.c.y * V0);
        T lv;
        const T d1 = dedfdn<UH>(a.ph, density(a.ph, phi3), e.it + 1, off);
        if (qf) lv = (T)0.5 * (V0 + e.vx1 + d1 + e.raux);
        else lv = (T)0.5 * (V0 + e.vx1 + d1);
        a.vbuf[off] = lv;
        return mul_exp(psi0, a.ph.c.x * lv, a.ph.c.y * lv);
    } else if constexpr (MM == MID_PART1) {
        // orginal/gpe_engine.cu:803-815 with the density sum of gpe_normalize fused
        const T V0 = pot2<UH>(a, e.vx, modify_at<UH>(a, e, e.aux, e.it), qf, e.raux, e.it, off);
        const cplx<T> phi3 = mul_exp(v, a.ph.c.x * V0, a.ph.c.y * V0);
        acc[0] += dens64((double)a.ph.gamma, (double)phi3.x, (double)phi3.y);
        return phi3;
    } else if constexpr (MM == MID_VSTEP3) {
        // orginal/gpe_engine.cu:870-878; the factor exp(c V1) is recomputed from V1
        const T lv = e.raux;
        cplx<T> psi1 = mul_exp(v, a.ph.c.x * lv, a.ph.c.y * lv);
        if (a.flags & F_RED_NORM) acc[0] += dens64((double)a.ph.gamma, (double)psi1.x, (double)psi1.y);
        if (a.flags & (F_NEXT_VSTEP1 | F_NEXT_PSI)) {
            if (a.flags & F_IMPRINT_NEXT) psi1 = apply_pin(a, psi1, e.c);
            a.psiA[off] = psi1;
            if (a.flags & F_NEXT_VSTEP1) {
                // the next step's Vstep1, at time index it + 1
                const cplx<T> m1 = modify_at<UH>(a, e, psi1, e.it + 1);
                const T l1 = pot2<UH>(a, e.vx1, m1, false, (T)0, e.it + 1, off);
                return mul_exp(m1, a.ph.c.x * l1, a.ph.c.y * l1);
            }
            return psi1;
        }
        a.psiA[off] = psi1;
        return psi1;
    } else if constexpr (MM == MID_QF_VSTEP1) {
        // V_qf = Im(conj psi0 * F^-1[qf' k^2 F psi0]) / (rho0 + eps) (orginal 911-922), then Vstep1_qf
        const cplx<T> psi0 = e.aux;
        const T r0 = density(a.ph, psi0);
        const T q = (psi0.x * v.y + (-psi0.y) * v.x) / (r0 + (T)1.0e-12);
        a.vq[off] = q;
        const cplx<T> m0 = modify_at<UH>(a, e, psi0, e.it);
        const T lv = pot2<UH>(a, e.vx, m0, true, q, e.it, off);
        return mul_exp(m0, a.ph.c.x * lv, a.ph.c.y * lv);
    } else if constexpr (MM == MID_KIN_SUM) {
        const double k2 = a.kx64[e.c.ix] * a.kx64[e.c.ix] + a.ky64[e.c.iy] * a.ky64[e.c.iy] +
                          a.kz64[e.c.iz] * a.kz64[e.c.iz];
        acc[0] += k2 * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        return v;
    } else if constexpr (MM == MID_VPP_SUM) {
        const double f = vpp_k(a.up64, a.kx64[e.c.ix], a.ky64[e.c.iy], a.kz64[e.c.iz]);
        acc[0] += f * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        return v;
    } else if constexpr (MM == MID_OVERLAP_RE) {
        const cplx<T> p = e.aux;
        a.rout[off] = (double)(p.x * v.x + p.y * v.y);
        return v;
    } else if constexpr (MM == MID_EXPT || MM == MID_EXPTC || MM == MID_K2 || MM == MID_KA || MM == MID_VPP ||
                         MM == MID_EXPT_SEP) {
        return mid_op_lean(a, e.c, v);
    } else {
        return v;
    }
}

template <bool UH, typename T>
__device__ __forceinline__ cplx<T> mid_op2(const PassArgs<T>& a, const ElemIn<T>& e, cplx<T> v, double* acc) {
    switch (a.mid_mode) {
    case MID_VSTEP2: return mid_op_m<MID_VSTEP2, UH>(a, e, v, acc);
    case MID_PART1: return mid_op_m<MID_PART1, UH>(a, e, v, acc);
    case MID_VSTEP3: return mid_op_m<MID_VSTEP3, UH>(a, e, v, acc);
    case MID_QF_VSTEP1: return mid_op_m<MID_QF_VSTEP1, UH>(a, e, v, acc);
    case MID_KIN_SUM: return mid_op_m<MID_KIN_SUM, UH>(a, e, v, acc);
    case MID_VPP_SUM: return mid_op_m<MID_VPP_SUM, UH>(a, e, v, acc);
    case MID_OVERLAP_RE: return mid_op_m<MID_OVERLAP_RE, UH>(a, e, v, acc);
    case MID_EXPT:
    case MID_EXPTC:
    case MID_K2:
    case MID_KA:
    case MID_VPP:
    case MID_EXPT_SEP: return mid_op_lean(a, e.c, v);
    default: return v;
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

// ------------------------------------------------------ the pass kernel
//
// Each transform of a pass runs its Stockham stages as
//   first stage   global memory -> butterfly registers -> LDS
//   middle stages LDS -> LDS
//   last stage    LDS -> butterfly registers
// (a single-stage length goes global -> registers directly).  A second
// transform in the other direction runs the reversed radix plan, whose first
// stage takes the same elements per thread as the last stage of the first
// transform, so the mid op and the hand-over stay in registers.  The final
// registers go straight to global memory.

// The alternative radix schedule of a length (launcher variant bit 6, chosen
// by the autotuner for the strided turn-around passes): the radices in the
// other order, or, where that is the same plan, the first radix split
// (64 = 4.16, 256 = 4.4.16, 512 = 4.8.16 for complex128); lengths of one
// stage have none.
constexpr FftPlan alt_plan(int N, bool c64) {
    FftPlan p = make_plan(N, c64);
    bool palindrome = true;
    for (int i = 0; i < p.ns / 2; i++) palindrome = palindrome && (p.r[i] == p.r[p.ns - 1 - i]);
    if (p.ns < 2) return p;
    if (!palindrome) {
        for (int i = 0; i < p.ns / 2; i++) {
            const int t = p.r[i];
            p.r[i] = p.r[p.ns - 1 - i];
            p.r[p.ns - 1 - i] = t;
        }
        return p;
    }
    FftPlan q{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    if (N == 64) { q.ns = 2; q.r[0] = 4; q.r[1] = 16; return q; }
    if (N == 256) { q.ns = 3; q.r[0] = 4; q.r[1] = 4; q.r[2] = 16; return q; }
    if (N == 512) { q.ns = 3; q.r[0] = 4; q.r[1] = 8; q.r[2] = 16; return q; }
    return p;
}

constexpr bool same_plan(const FftPlan& a, const FftPlan& b) {
    if (a.ns != b.ns) return false;
    for (int i = 0; i < a.ns; i++)
        if (a.r[i] != b.r[i]) return false;
    return true;
}

// whether a length has an alternative schedule
constexpr bool has_alt_plan(int N, bool c64) { return !same_plan(alt_plan(N, c64), make_plan(N, c64)); }

// plan of a transform: PV 0 the default schedule, 1 the alternative; rev
// reverses the radix order (the second transform of a turn-around pass)
constexpr FftPlan plan_of(int N, bool rev, bool c64, int pv = 0) {
    FftPlan p = pv ? alt_plan(N, c64) : make_plan(N, c64);
    if (rev)
        for (int i = 0; i < p.ns / 2; i++) {
            int t = p.r[i];
            p.r[i] = p.r[p.ns - 1 - i];
            p.r[p.ns - 1 - i] = t;
        }
    return p;
}

template <typename T>
__device__ __forceinline__ unsigned axis_stride(const PassArgs<T>& a) {
    return a.axis == 2 ? 1u : a.axis == 1 ? (unsigned)a.ph.nz : (unsigned)(a.ph.ny * a.ph.nz);
}

// pencil slot -> global element offsets and lattice coordinates
struct Pencil {
    unsigned base;    // offset of element 0
    unsigned stride;  // element stride along the axis
    unsigned c0, c1;  // the two fixed coordinates
};

// AX >= 0: the axis known at compile time (no branch), -1: read from the arguments
template <int AX = -1, typename T>
__device__ __forceinline__ Pencil make_pencil(const PassArgs<T>& a, int N, unsigned p) {
    Pencil q;
    const unsigned nx = (unsigned)a.ph.nx, ny = (unsigned)a.ph.ny, nz = (unsigned)a.ph.nz;
    const int axis = AX >= 0 ? AX : a.axis;
    if (axis == 2) {  // z pencils (b, ix, iy) in order: p = (b nx + ix) ny + iy
        const unsigned bx = p / ny;
        q.c1 = p - bx * ny;
        q.c0 = bx - (bx / nx) * nx;
        q.base = p * (unsigned)N;
        q.stride = 1;
    } else if (axis == 1) {  // y pencils (b, ix, iz): p = (b nx + ix) nz + iz
        const unsigned bx = p / nz;
        q.c1 = p - bx * nz;
        q.c0 = bx - (bx / nx) * nx;
        q.base = bx * (ny * nz) + q.c1;
        q.stride = nz;
    } else {  // x pencils (b, iy, iz): p = b ny nz + iy nz + iz
        const unsigned plane = ny * nz;
        const unsigned b = p / plane;
        const unsigned pl = p - b * plane;
        q.c0 = pl / nz;
        q.c1 = pl - q.c0 * nz;
        q.base = b * a.n1 + pl;
        q.stride = plane;
    }
    return q;
}

// selects on the (uniform) axis, no branches: a branch per point would split
// the unrolled per-point code into basic blocks the scheduler cannot interleave
template <int AX = -1, typename T>
__device__ __forceinline__ Coord elem(const PassArgs<T>& a, const Pencil& q, unsigned i) {
    Coord c;
    c.off = q.base + i * q.stride;
    const int axis = AX >= 0 ? AX : a.axis;
    const bool ax0 = (axis == 0), ax1 = (axis == 1);
    c.ix = ax0 ? i : q.c0;
    c.iy = ax0 ? q.c0 : (ax1 ? i : q.c1);
    c.iz = (ax0 || ax1) ? q.c1 : i;
    return c;
}

// butterfly b of a stage with N/R = M butterflies per pencil -> (pencil slot, j)
template <int M, int W, bool CONTIG>
__device__ __forceinline__ void bmap(int b, int& w, int& j) {
    if constexpr (CONTIG) { w = b / M; j = b % M; }
    else { j = b / W; w = b % W; }
}

// first stage from global memory (load op applied per element) into LDS
template <typename T, int N, int W, bool CONTIG, bool FULL, int S, int R, int AX = -1>
__device__ __forceinline__ void first_stage_global(const PassArgs<T>& a, unsigned p0, cplx<T>* lds, double* acc) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
    constexpr int PER = (NB + kThreads - 1) / kThreads;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        if ((NB % kThreads) == 0 || b < NB) {
            int w, j;
            bmap<M, W, CONTIG>(b, w, j);
            const unsigned p = p0 + (unsigned)w;
            cplx<T> v[R];
            if (p < a.M) {
                const Pencil pc = make_pencil<AX>(a, N, p);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const Coord c = elem<AX>(a, pc, (unsigned)(j + r * M));
                    v[r] = a.src[c.off];
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) v[r] = mk((T)0, (T)0);
            }
            dft<R, S>(v);
#pragma unroll
            for (int r = 0; r < R; r++) lds[lidx<T, N, W, CONTIG>(w, j * R + r)] = v[r];
        }
    }
}

// first stage of a lean pass from HBM into the LDS tile, calling pre() once
// after the first loads are issued (the workgroup's table staging then runs
// while the tile's rows are in flight)
template <typename T, int N, int W, bool CONTIG, int S, int R, int AX, typename F>
__device__ __forceinline__ void first_stage_global_pre(const PassArgs<T>& a, unsigned p0, cplx<T>* lds, F&& pre) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
    constexpr int PER = (NB + kThreads - 1) / kThreads;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        int w = 0, j = 0;
        cplx<T> v[R];
        const bool in = (NB % kThreads) == 0 || b < NB;
        if (in) {
            bmap<M, W, CONTIG>(b, w, j);
            const unsigned p = p0 + (unsigned)w;
            if (p < a.M) {
                const Pencil pc = make_pencil<AX>(a, N, p);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const Coord c = elem<AX>(a, pc, (unsigned)(j + r * M));
                    v[r] = a.src[c.off];
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) v[r] = mk((T)0, (T)0);
            }
        }
        if (q == 0) pre();
        if (in) {
            dft<R, S>(v);
#pragma unroll
            for (int r = 0; r < R; r++) lds[lidx<T, N, W, CONTIG>(w, j * R + r)] = v[r];
        }
    }
}

// first stage from registers already holding (w, j + r N/R) into LDS
template <typename T, int N, int W, bool CONTIG, int S, int R, int PER>
__device__ __forceinline__ void first_stage_regs(cplx<T> (&v)[PER][R], cplx<T>* lds) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        if ((NB % kThreads) == 0 || b < NB) {
            int w, j;
            bmap<M, W, CONTIG>(b, w, j);
            dft<R, S>(v[q]);
#pragma unroll
            for (int r = 0; r < R; r++) lds[lidx<T, N, W, CONTIG>(w, j * R + r)] = v[q][r];
        }
    }
}

// stages [STAGE, LAST) of plan (N, REV, PV) in LDS
template <typename T, int N, int W, bool CONTIG, int S, bool REV, int STAGE, int LAST, int PV = 0>
__device__ __forceinline__ void lds_stages(cplx<T>* lds, const TwE<T>* tw) {
    if constexpr (STAGE < LAST) {
        constexpr FftPlan P = plan_of(N, REV, sizeof(T) == 4, PV);
        fft_stage<T, N, W, CONTIG, S, P.r[STAGE], plan_span(P, STAGE)>(lds, tw);
        lds_stages<T, N, W, CONTIG, S, REV, STAGE + 1, LAST, PV>(lds, tw);
    }
}

// last stage (span N/R) from LDS into registers: v[q][r] is element (w, j + r N/R)
template <typename T, int N, int W, bool CONTIG, int S, int R, int PER>
__device__ __forceinline__ void last_stage_regs(const cplx<T>* lds, const TwE<T>* tw,
                                                cplx<T> (&v)[PER][R]) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        if ((NB % kThreads) == 0 || b < NB) {
            int w, j;
            bmap<M, W, CONTIG>(b, w, j);
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = lds[lidx<T, N, W, CONTIG>(w, j + r * M)];
#pragma unroll
            for (int r = 1; r < R; r++) v[q][r] = tw_mul<S>(v[q][r], tw[j * r]);  // twiddle w_N^(j r)
            dft<R, S>(v[q]);
        }
    }
}

// registers (w, j + r N/R) <- global memory with the load op (single-stage lengths)
template <typename T, int N, int W, bool CONTIG, bool FULL, int R, int PER, int AX = -1>
__device__ __forceinline__ void load_regs(const PassArgs<T>& a, unsigned p0, cplx<T> (&v)[PER][R], double* acc) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        int w, j;
        bmap<M, W, CONTIG>(b, w, j);
        const unsigned p = p0 + (unsigned)w;
        const bool ok = ((NB % kThreads) == 0 || b < NB) && p < a.M;
        const Pencil pc = make_pencil<AX>(a, N, ok ? p : 0u);
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (ok) {
                const Coord c = elem<AX>(a, pc, (unsigned)(j + r * M));
                v[q][r] = a.src[c.off];
            } else {
                v[q][r] = mk((T)0, (T)0);
            }
        }
    }
}

template <int MK, bool CJ, typename T, int N, int W, bool CONTIG, int R, int PER, int AX = -1>
__device__ __forceinline__ void mid_regs_k(const PassArgs<T>& a, unsigned p0, cplx<T> (&v)[PER][R],
                                           const float4* kx = nullptr) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
    // keep the factor-table loads here: hoisted above the first transform they
    // would hold registers through it (a compiler-only barrier, no instruction)
    asm volatile("" ::: "memory");
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        int w, j;
        bmap<M, W, CONTIG>(b, w, j);
        const unsigned p = p0 + (unsigned)w;
        if (((NB % kThreads) == 0 || b < NB) && p < a.M) {
            const Pencil pc = make_pencil<AX>(a, N, p);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const Coord c = elem<AX>(a, pc, (unsigned)(j + r * M));
                v[q][r] = mid_op_lean_k<MK, CJ, T, AX>(a, c, v[q][r], kx);
            }
        }
    }
}

template <typename T, int N, int W, bool CONTIG, int R, int PER, int AX = -1>
__device__ __forceinline__ void store_regs(const PassArgs<T>& a, unsigned p0, const cplx<T> (&v)[PER][R]) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        int w, j;
        bmap<M, W, CONTIG>(b, w, j);
        const unsigned p = p0 + (unsigned)w;
        if (((NB % kThreads) == 0 || b < NB) && p < a.M) {
            const Pencil pc = make_pencil<AX>(a, N, p);
#pragma unroll
            for (int r = 0; r < R; r++) a.dst[pc.base + (unsigned)(j + r * M) * pc.stride] = v[q][r];
        }
    }
}

// pass shapes: one transform forward / inverse, or two with the mid op between
enum PassShape { SHAPE_F = 0, SHAPE_I = 1, SHAPE_FI = 2, SHAPE_IF = 3 };

// Lean passes (plain load and store, at most a table factor in the middle):
// everything between the first load and the last store stays in registers or LDS.
// tile width of a lean pass; WD = 2 halves it (an autotuner choice for strided passes)
template <typename T, int N, int WD>
constexpr int lean_w() {
    return tile_w<T, N>() / WD > 0 ? tile_w<T, N>() / WD : 1;
}

// What follows the first stage of a lean pass of two or more stages (the first
// stage's output is in LDS): the remaining stages, the mid op, the optional
// second transform and the store of tile p0.
template <typename T, int N, int W, bool CONTIG, int S1, int S2, int MK, bool CJ, int AX, int PV = 0>
__device__ __forceinline__ void lean_tail(const PassArgs<T>& a, unsigned p0, cplx<T>* lds, const TwE<T>* tw,
                                          const float4* kx = nullptr) {
    constexpr FftPlan P = plan_of(N, false, sizeof(T) == 4, PV);
    constexpr int NSG = P.ns;
    constexpr int RL = P.r[NSG - 1];
    constexpr int PERL = (W * (N / RL) + kThreads - 1) / kThreads;
    cplx<T> v[PERL][RL];
    lds_stages<T, N, W, CONTIG, S1, false, 1, NSG - 1, PV>(lds, tw);
    last_stage_regs<T, N, W, CONTIG, S1, RL, PERL>(lds, tw, v);
    if constexpr (MK >= 0) mid_regs_k<MK, CJ, T, N, W, CONTIG, RL, PERL, AX>(a, p0, v, kx);
    if constexpr (S2 != 0) {
        constexpr FftPlan Q = plan_of(N, true, sizeof(T) == 4, PV);
        static_assert(Q.r[0] == RL, "reversed plan must start with the last radix");
        constexpr int RF = Q.r[NSG - 1];
        constexpr int PERF = (W * (N / RF) + kThreads - 1) / kThreads;
        __syncthreads();  // the first transform's last-stage reads are done
        first_stage_regs<T, N, W, CONTIG, S2, RL, PERL>(v, lds);
        __syncthreads();
        lds_stages<T, N, W, CONTIG, S2, true, 1, NSG - 1, PV>(lds, tw);
        cplx<T> u[PERF][RF];
        last_stage_regs<T, N, W, CONTIG, S2, RF, PERF>(lds, tw, u);
        store_regs<T, N, W, CONTIG, RF, PERF, AX>(a, p0, u);
    } else {
        store_regs<T, N, W, CONTIG, RL, PERL, AX>(a, p0, v);
    }
}

// Tile of workgroup b.  Strided tiles whose rows are shorter than a 128-byte
// line share each line with the next G - 1 tiles.  Workgroups are dispatched
// round-robin over the 8 XCDs, each with its own L2, so those G tiles go to
// workgroups b, b + 8, ..., b + 8 (G - 1) of one XCD rather than to
// neighbouring workgroups on different XCDs (a trailing partial group keeps
// the plain order).
template <typename T, int W, bool CONTIG>
__device__ __forceinline__ unsigned tile_of_block(unsigned b) {
    constexpr int RB = W * (int)sizeof(cplx<T>);
    if constexpr (CONTIG || RB >= 128) {
        return b;
    } else {
        constexpr unsigned G = 128 / RB, X = 8, GX = G * X;
        const unsigned q = b / GX, r = b - q * GX;
        if ((q + 1) * GX > gridDim.x) return b;
        return q * GX + (r % X) * G + r / X;
    }
}

// MK: compile-time kind of the mid op (MidKind), -1 for none; CJ conjugates it
// stage(): writes the workgroup's LDS tables (twiddles, factor), called
// once the tile's first loads are issued; the first barrier after it orders
// the writes before any use
template <typename T, int N, bool CONTIG, int S1, int S2, int WD, int MK, bool CJ, int AX, int PV, typename F>
__device__ __forceinline__ void lean_body(const PassArgs<T>& a, cplx<T>* lds, const TwE<T>* tw, const float4* kx,
                                          F&& stage) {
    constexpr int W = lean_w<T, N, WD>();
    constexpr FftPlan P = plan_of(N, false, sizeof(T) == 4, PV);
    constexpr int NSG = P.ns;
    constexpr int RL = P.r[NSG - 1];
    constexpr int PERL = (W * (N / RL) + kThreads - 1) / kThreads;
    const unsigned p0 = tile_of_block<T, W, CONTIG>(blockIdx.x) * (unsigned)W;
    if constexpr (NSG == 1) {
        cplx<T> v[PERL][RL];
        load_regs<T, N, W, CONTIG, false, RL, PERL, AX>(a, p0, v, nullptr);
        stage();
#pragma unroll
        for (int q = 0; q < PERL; q++) dft<RL, S1>(v[q]);
        if (kx) __syncthreads();  // the LDS factor table is written
        if constexpr (MK >= 0) mid_regs_k<MK, CJ, T, N, W, CONTIG, RL, PERL, AX>(a, p0, v, kx);
        if constexpr (S2 != 0) {
#pragma unroll
            for (int q = 0; q < PERL; q++) dft<RL, S2>(v[q]);
        }
        store_regs<T, N, W, CONTIG, RL, PERL, AX>(a, p0, v);
    } else {
        // strided tiles: the waves issuing a tile's loads (and its first
        // stage) run at raised priority, so their requests go out ahead of
        // the transforms of the CU's other workgroups (256^3 complex64: y pass
        // 51.8 -> 50.8 us, x 50.2 -> 49.5 us; the same in the z-wave passes
        // measured slower)
        if constexpr (!CONTIG) __builtin_amdgcn_s_setprio(3);
        first_stage_global_pre<T, N, W, CONTIG, S1, P.r[0], AX>(a, p0, lds, stage);
        if constexpr (!CONTIG) __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        lean_tail<T, N, W, CONTIG, S1, S2, MK, CJ, AX, PV>(a, p0, lds, tw, kx);
    }
}

// ------------------------------------------------ pipelined strided passes
//
// complex64 lean passes along a strided axis (y or x) in persistent
// workgroups: workgroup b transforms tiles b, b + gridDim.x, ...  The rows of
// the next tile are copied from HBM straight into an LDS staging buffer
// (global_load_lds_dwordx4: 64 lanes x 16 bytes, 8 rows of 16 columns per
// instruction, no registers) as soon as the current tile's first stage has
// read its own rows, so the next tile's loads are in flight while the
// current tile's remaining stages, mid op, second transform and stores run.
// Every row of a tile is W contiguous elements at base + i * stride (the
// launcher checks the lattice allows it).

// lengths with a pipelined kernel: two or more stages, a tile of at most 4096
// elements (32 KB of staging) and whole 16-byte lane copies per row
template <int N>
constexpr bool pipe_ok() {
    constexpr int W = lean_w<float, N, 1>();
    if (W < 2 || W > 128 || W % 2 != 0 || N * W > 4096 || plan_of(N, false, true).ns < 2) return false;
    return N % ((128 / W) * (kThreads / 64)) == 0;  // whole wave instructions of 128 / W rows per wave
}

// issues the copies of tile p0's N rows into stage ([row][W], linear)
template <int N, int W, int AX>
__device__ __forceinline__ void pipe_fetch(const PassArgs<float>& a, unsigned p0, cplx<float>* stage) {
    constexpr int LPR = W / 2;          // lanes per row, 16 bytes (2 elements) each
    constexpr int RPI = 64 / LPR;       // rows per wave instruction
    constexpr int NI = N / RPI;         // instructions per tile
    constexpr int PERW = NI / (kThreads / 64);
    static_assert(PERW * (kThreads / 64) * RPI == N, "tile rows must split evenly over the waves");
    const Pencil pc = make_pencil<AX>(a, N, p0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned rr = (unsigned)(lane / LPR), cc = 2u * (unsigned)(lane % LPR);
#pragma unroll
    for (int q = 0; q < PERW; q++) {
        const int g = wv * PERW + q;
        const cplx<float>* src = a.src + pc.base + ((unsigned)(g * RPI) + rr) * pc.stride + cc;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (__attribute__((address_space(3))) void*)(stage + g * 2 * 64), 16, 0, 0);
    }
}

// the first Stockham stage from the staging buffer into the tile
template <int N, int W, int S, int R>
__device__ __forceinline__ void first_stage_stage(const cplx<float>* stage, cplx<float>* lds) {
    constexpr int M = N / R;
    constexpr int NB = W * M;
    constexpr int PER = (NB + kThreads - 1) / kThreads;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kThreads;
        if ((NB % kThreads) == 0 || b < NB) {
            int w, j;
            bmap<M, W, false>(b, w, j);
            cplx<float> v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = stage[(j + r * M) * W + w];
            dft<R, S>(v);
#pragma unroll
            for (int r = 0; r < R; r++) lds[lidx<float, N, W, false>(w, j * R + r)] = v[r];
        }
    }
}

template <int N, int SHAPE, int MK, bool CJ, int AX>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
void fft_pipe_kernel(const PassArgs<float> a) {
    constexpr int W = lean_w<float, N, 1>();
    constexpr FftPlan P = plan_of(N, false, true);
    constexpr int S1 = (SHAPE == SHAPE_F || SHAPE == SHAPE_FI) ? -1 : 1;
    constexpr int S2 = (SHAPE == SHAPE_FI) ? 1 : (SHAPE == SHAPE_IF) ? -1 : 0;
    __shared__ TwE<float> tw[N];
    __shared__ cplx<float> stage[W * N];
    __shared__ cplx<float> lds[lds_elems<float, N, W, false>()];
    for (int i = threadIdx.x; i < N; i += kThreads) {
        const cplx<float> h = a.tw[i], lo = a.twlo[i];
        tw[i].t = make_float4(h.x, h.y, lo.x, lo.y);
    }
    // the x pass's K_x pairs in LDS, as in the lean passes (ordered before
    // their first use by the loop's first barrier)
    constexpr bool KXL = MK == MK_SEP && AX == 0;
    __shared__ float4 kxs[KXL ? N : 1];
    if constexpr (KXL) {
        const float4* src = reinterpret_cast<const float4*>(a.kx_hl);
        for (int i = threadIdx.x; i < N; i += kThreads) kxs[i] = src[i];
    }
    const unsigned ntiles = a.M / (unsigned)W;
    unsigned t = blockIdx.x;
    if (t < ntiles) pipe_fetch<N, W, AX>(a, t * (unsigned)W, stage);
    for (; t < ntiles; t += gridDim.x) {
        // this tile's rows have landed in the staging buffer (the wait also
        // covers the previous tile's stores), for every wave of the group
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        first_stage_stage<N, W, S1, P.r[0]>(stage, lds);
        __syncthreads();  // the staging buffer is free, the tile is in LDS
        if (t + gridDim.x < ntiles) pipe_fetch<N, W, AX>(a, (t + gridDim.x) * (unsigned)W, stage);
        lean_tail<float, N, W, false, S1, S2, MK, CJ, AX>(a, t * (unsigned)W, lds, tw, KXL ? kxs : nullptr);
    }
}

// Physics passes.  Every per-element global input is gathered into LDS up
// front by unrolled plain loads (the FFT input; psi0 as a complex side tile;
// V_qf or V1 as a real side tile), together with per-pencil coordinates and the
// separable external potential (V_xy per pencil, V_z per point).  The load and
// mid ops then read only LDS, element by element in rolled loops (the op code
// is not replicated per register), and their global outputs are plain stores.
// The transforms run in LDS, the last stage into registers for the store.

template <typename T, int N, bool CONTIG>
constexpr int full_tile_w() {
    int w = tile_w<T, N>();
    if (CONTIG && w > 1) w /= 2;  // three tiles of LDS per workgroup
    // complex128 128 and 256: the 2048-element tiles (the wider lean tiles
    // measured slower here: 137 -> 197 us for the energy's kinetic pass)
    if (sizeof(T) == 8 && (N == 128 || N == 256) && w > 1) w /= 2;
    // complex128 512 along z (no wave split there): one pencil per workgroup,
    // 34 KB of LDS, so that four workgroups share a CU at the four-wave bound
    // of fft_pass_kernel (512^3: 70.5 -> 87.6 steps/s)
    if (sizeof(T) == 8 && N == 512 && CONTIG) w = 1;
    return w;
}

template <typename T, int N, bool CONTIG>
struct FullLds {
    static constexpr int W = full_tile_w<T, N, CONTIG>();
    static constexpr int TILE = lds_elems<T, N, W, CONTIG>();
    static constexpr int SIDE = CONTIG ? TILE : 1;   // side tiles only in z passes
    cplx<T> main[TILE];
    cplx<T> aux[SIDE];
    T raux[SIDE];
    T vz[CONTIG ? N : 1];
    T vxy[W];
    unsigned base[W], c0[W], c1[W];
    bool ok[W];
};

template <typename T>
__device__ __forceinline__ bool needs_psi0(const PassArgs<T>& a) {
    return a.load_mode == LOAD_PART2 || a.mid_mode == MID_VSTEP2 || a.mid_mode == MID_PART1 ||
           a.mid_mode == MID_QF_VSTEP1 || a.mid_mode == MID_OVERLAP_RE;
}
template <typename T>
__device__ __forceinline__ const T* real_side(const PassArgs<T>& a) {
    if (a.mid_mode == MID_VSTEP3) return a.vbuf;
    const bool qf = (a.flags & F_QF) != 0;
    if (qf && (a.load_mode == LOAD_PART2 || a.mid_mode == MID_VSTEP2 || a.mid_mode == MID_PART1 ||
               (a.load_mode == LOAD_PSI0 && (a.flags & F_VSTEP1))))
        return a.vq;
    return nullptr;
}

template <bool UH, typename T, int N, bool CONTIG>
__device__ __forceinline__ ElemIn<T> elem_in(const PassArgs<T>& a, const FullLds<T, N, CONTIG>& L, int w, int i,
                                             bool sep, T vscale, unsigned it) {
    ElemIn<T> e;
    e.it = it;
    const unsigned ui = (unsigned)i;
    if (a.axis == 2) { e.c.ix = L.c0[w]; e.c.iy = L.c1[w]; e.c.iz = ui; e.c.off = L.base[w] + ui; }
    else if (a.axis == 1) { e.c.ix = L.c0[w]; e.c.iy = ui; e.c.iz = L.c1[w]; e.c.off = L.base[w] + ui * (unsigned)a.ph.nz; }
    else { e.c.ix = ui; e.c.iy = L.c0[w]; e.c.iz = L.c1[w]; e.c.off = L.base[w] + ui * (unsigned)(a.ph.ny * a.ph.nz); }
    if constexpr (CONTIG) {
        e.vx = sep ? vscale * (L.vxy[w] + L.vz[i]) : vext<UH>(a.ph, e.c.ix, e.c.iy, e.c.iz, e.c.off, it);
        e.vx1 = e.vx;
        if constexpr (UH) {
            if (a.ph.vext == VEXT_USER) e.vx1 = vext<UH>(a.ph, e.c.ix, e.c.iy, e.c.iz, e.c.off, it + 1);
        }
        const int li = lidx<T, N, FullLds<T, N, CONTIG>::W, CONTIG>(w, i);
        e.aux = L.aux[li];
        e.raux = L.raux[li];
    } else {
        e.vx = (T)0;
        e.vx1 = (T)0;
        e.aux = mk((T)0, (T)0);
        e.raux = (T)0;
    }
    return e;
}

template <typename T, int N, bool CONTIG, int S1, int S2, bool UH>
__device__ __forceinline__ void full_body(const PassArgs<T>& a, FullLds<T, N, CONTIG>& L, const TwE<T>* tw, double* acc) {
    constexpr int W = FullLds<T, N, CONTIG>::W;
    constexpr int TILE = W * N;
    constexpr int PERE = (TILE + kThreads - 1) / kThreads;
    constexpr FftPlan P = plan_of(N, false, sizeof(T) == 4);
    constexpr int NSG = P.ns;
    constexpr int RL = P.r[NSG - 1];
    constexpr int PERL = (W * (N / RL) + kThreads - 1) / kThreads;
    const unsigned p0 = blockIdx.x * (unsigned)W;
    const unsigned it = time_index(a.ph);

    // per-pencil coordinates and the pencil part of a separable potential
    const int vk = a.ph.vext;
    const bool sep = CONTIG && (vk == VEXT_HARMONIC || vk == VEXT_COSINE || vk == VEXT_WOODS_SAXON);
    const T vscale = (vk == VEXT_COSINE) ? a.ph.up[0] : (T)1;
    if (threadIdx.x < W) {
        const int w = threadIdx.x;
        const unsigned p = p0 + (unsigned)w;
        L.ok[w] = p < a.M;
        const Pencil pc = make_pencil(a, N, L.ok[w] ? p : 0u);
        L.base[w] = pc.base;
        L.c0[w] = pc.c0;
        L.c1[w] = pc.c1;
        if (sep) {
            const int nx = a.ph.nx, ny = a.ph.ny;
            if (vk == VEXT_WOODS_SAXON) L.vxy[w] = vext(a.ph, pc.c0, pc.c1, 0u, 0u, 0u);
            else L.vxy[w] = a.ph.vtab[pc.c0] + a.ph.vtab[nx + pc.c1];
            (void)ny;
        }
    }
    if constexpr (CONTIG) {
        if (sep)
            for (int i = threadIdx.x; i < N; i += kThreads)
                L.vz[i] = (vk == VEXT_WOODS_SAXON) ? (T)0 : a.ph.vtab[a.ph.nx + a.ph.ny + i];
    }

    // gather: unrolled plain loads of every input of the tile
    const cplx<T>* msrc = (a.load_mode == LOAD_PSI0) ? a.psiA : a.src;
    const bool want_psi0 = CONTIG && needs_psi0(a);
    const T* rsrc = CONTIG ? real_side(a) : nullptr;
#pragma unroll
    for (int q = 0; q < PERE; q++) {
        const int e = threadIdx.x + q * kThreads;
        if ((TILE % kThreads) == 0 || e < TILE) {
            int w, i;
            if constexpr (CONTIG) { w = e / N; i = e % N; }
            else { i = e / W; w = e % W; }
            const unsigned p = p0 + (unsigned)w;
            const int li = lidx<T, N, W, CONTIG>(w, i);
            if (p < a.M) {
                const Pencil pc = make_pencil(a, N, p);
                const unsigned off = pc.base + (unsigned)i * pc.stride;
                L.main[li] = msrc[off];
                if constexpr (CONTIG) {
                    if (want_psi0) L.aux[li] = a.psiA[off];
                    if (rsrc) L.raux[li] = rsrc[off];
                }
            } else {
                L.main[li] = mk((T)0, (T)0);
            }
        }
    }
    __syncthreads();

    // load op, element by element from LDS
    if (a.load_mode != LOAD_PLAIN || (a.flags & F_RED_POT)) {
        for (int q = 0; q < PERE; q++) {
            const int e = threadIdx.x + q * kThreads;
            if ((TILE % kThreads) == 0 || e < TILE) {
                int w, i;
                if constexpr (CONTIG) { w = e / N; i = e % N; }
                else { i = e / W; w = e % W; }
                if (L.ok[w]) {
                    const int li = lidx<T, N, W, CONTIG>(w, i);
                    L.main[li] = load_op2<UH>(a, elem_in<UH>(a, L, w, i, sep, vscale, it), L.main[li], acc);
                }
            }
        }
        __syncthreads();
    }

    cplx<T> v[PERL][RL];
    if (a.mid_mode == MID_NONE) {
        lds_stages<T, N, W, CONTIG, S1, false, 0, NSG - 1>(L.main, tw);
        last_stage_regs<T, N, W, CONTIG, S1, RL, PERL>(L.main, tw, v);
        if (a.store_mode == STORE_PLAIN) store_regs<T, N, W, CONTIG, RL, PERL>(a, p0, v);
        return;
    }
    lds_stages<T, N, W, CONTIG, S1, false, 0, NSG>(L.main, tw);
    for (int q = 0; q < PERE; q++) {  // mid op, natural order
        const int e = threadIdx.x + q * kThreads;
        if ((TILE % kThreads) == 0 || e < TILE) {
            int w, i;
            if constexpr (CONTIG) { w = e / N; i = e % N; }
            else { i = e / W; w = e % W; }
            if (L.ok[w]) {
                const int li = lidx<T, N, W, CONTIG>(w, i);
                L.main[li] = mid_op2<UH>(a, elem_in<UH>(a, L, w, i, sep, vscale, it), L.main[li], acc);
            }
        }
    }
    __syncthreads();
    if constexpr (S2 != 0) {
        lds_stages<T, N, W, CONTIG, S2, false, 0, NSG - 1>(L.main, tw);
        last_stage_regs<T, N, W, CONTIG, S2, RL, PERL>(L.main, tw, v);
        if (a.store_mode == STORE_PLAIN) store_regs<T, N, W, CONTIG, RL, PERL>(a, p0, v);
    } else if (a.store_mode == STORE_PLAIN) {
        for (int q = 0; q < PERE; q++) {
            const int e = threadIdx.x + q * kThreads;
            if ((TILE % kThreads) == 0 || e < TILE) {
                int w, i;
                if constexpr (CONTIG) { w = e / N; i = e % N; }
                else { i = e / W; w = e % W; }
                if (L.ok[w]) a.dst[L.base[w] + (unsigned)i * axis_stride(a)] = L.main[lidx<T, N, W, CONTIG>(w, i)];
            }
        }
    }
}

// complex64 at 4 waves per SIMD, complex128 at 2, except the complex128
// 512-point z physics passes of the built-in physics: bounded to 4 waves they
// need 75-77 VGPRs and no spills (at 2 the compiler spends ~190); with user
// hooks they would spill, and at 96 and 192 points (wider tiles) too
template <typename T, int N, bool CONTIG, bool FULL, int SHAPE, int WD = 1, int MK = -1, bool CJ = false, int AX = -1,
          bool UH = false, int PV = 0>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : ((FULL && CONTIG && N == 512 && !UH) ? 4 : 2))))
void fft_pass_kernel(const PassArgs<T> a) {
    __shared__ TwE<T> tw[N];
    auto stage_tw = [&]() __attribute__((always_inline)) {
        for (int i = threadIdx.x; i < N; i += kThreads) {
            if constexpr (sizeof(T) == 4) {
                const cplx<float> h = a.tw[i], lo = a.twlo[i];
                tw[i].t = make_float4(h.x, h.y, lo.x, lo.y);
            } else {
                tw[i].t = a.tw[i];
            }
        }
    };
    // (the first barrier of every body orders these table writes before any use;
    // a single-stage length reads no table)
    constexpr int S1 = (SHAPE == SHAPE_F || SHAPE == SHAPE_FI) ? -1 : 1;
    constexpr int S2 = (SHAPE == SHAPE_FI) ? 1 : (SHAPE == SHAPE_IF) ? -1 : 0;
    if constexpr (!FULL) {
        constexpr int W = lean_w<T, N, WD>();
        __shared__ cplx<T> lds[lds_elems<T, N, W, CONTIG>()];
        // the x turn-around's K_x pairs in LDS: read from global memory per
        // point, their 16-byte loads cost the pass ~15 us of its ~66 at 256^3
        // complex64 (measured).  complex128 stages K_x or K_y (16 bytes per
        // entry) up to 256 points, where the extra LDS costs no workgroup per
        // CU: 256^3 y pass 110.5 -> 105.0 us, x 97.7 -> 95.1 us.  complex64's
        // y pass keeps its per-(iy, iz) table (K_y in LDS with a second
        // product measured slower)
        constexpr bool KXL = MK == MK_SEP && !CONTIG &&
                             (sizeof(T) == 4 ? AX == 0 : ((AX == 0 || AX == 1) && N <= 256));
        __shared__ float4 kxs[KXL ? N : 1];
        auto stage = [&]() __attribute__((always_inline)) {
            stage_tw();
            if constexpr (KXL) {
                const float4* src = reinterpret_cast<const float4*>(
                    sizeof(T) == 4 ? (const void*)a.kx_hl : (AX == 0 ? (const void*)a.Kx : (const void*)a.Ky));
                for (int i = threadIdx.x; i < N; i += kThreads) kxs[i] = src[i];
            }
        };
        // complex128 stages the tables once the tile's first loads are in
        // flight (256^3: y pass 105.3 -> 103.5 us, x 95.3 -> 94.2 us);
        // complex64 stages them first (deferred, its y pass measured 50.9 ->
        // 53.7 us)
        if constexpr (sizeof(T) == 4) {
            stage();
            lean_body<T, N, CONTIG, S1, S2, WD, MK, CJ, (CONTIG ? 2 : AX), PV>(a, lds, tw, KXL ? kxs : nullptr,
                                                                               []() {});
        } else {
            lean_body<T, N, CONTIG, S1, S2, WD, MK, CJ, (CONTIG ? 2 : AX), PV>(a, lds, tw, KXL ? kxs : nullptr, stage);
        }
    } else {
        stage_tw();
        __shared__ FullLds<T, N, CONTIG> L;
        __shared__ double red[kThreads / 64];
        double acc[3] = {0.0, 0.0, 0.0};
        full_body<T, N, CONTIG, S1, S2, UH>(a, L, tw, acc);
        // FP64 partial sums of this workgroup (indices fixed at compile time)
        if (a.nred > 0) {
            const double s0 = block_sum(acc[0], red);
            if (threadIdx.x == 0) a.partials[blockIdx.x] = s0;
        }
        if (a.nred > 1) {
            const double s1 = block_sum(acc[1], red);
            if (threadIdx.x == 0) a.partials[(size_t)gridDim.x + blockIdx.x] = s1;
        }
        if (a.nred > 2) {
            const double s2 = block_sum(acc[2], red);
            if (threadIdx.x == 0) a.partials[2 * (size_t)gridDim.x + blockIdx.x] = s2;
        }
    }
}

// the pass can run in the lean kernel
template <typename T>
inline bool pass_is_lean(const PassArgs<T>& a) {
    return a.load_mode == LOAD_PLAIN && a.flags == 0 && a.nred == 0 &&
           (a.mid_mode == MID_NONE || a.mid_mode == MID_EXPT || a.mid_mode == MID_EXPTC || a.mid_mode == MID_K2 ||
            a.mid_mode == MID_KA || a.mid_mode == MID_VPP || a.mid_mode == MID_EXPT_SEP);
}

inline int pass_shape(int dir1, int dir2) {
    if (dir2 == 0) return dir1 < 0 ? SHAPE_F : SHAPE_I;
    return dir1 < 0 ? SHAPE_FI : SHAPE_IF;
}

}  // namespace gpe
