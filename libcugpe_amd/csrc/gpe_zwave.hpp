// z-axis passes with one wave64 per group of pencils, registers end to end.
//
// The z axis is the contiguous one, so a pencil of N points is one run of
// N elements.  A wave64 takes 64/A pencils with A lanes each; lane l of a
// pencil holds the B = N/A elements n = l + A b (b = 0..B-1) in registers.
// With n = l + A b and k = kb + B ka,
//
//   X[kb + B ka] = sum_l w_A^(l ka) w_N^(l kb) sum_b x[l + A b] w_B^(b kb)
//
// so one transform is: a B-point DFT in each lane, the twiddle w_N^(l kb), an
// exchange through the wave's own LDS rows, and A-point DFTs over the lanes.
// After the exchange lane l owns the columns kb = l + A m (m < B/A), and the
// output X[l + A m + B ka] lands in register m + (B/A) ka: index l + A b'
// again (B is a multiple of A; at 512 points, B = 2A = 32, the exchange runs
// in two phases of A rows to keep the LDS area of a 256-point pencil).  Every transform maps this layout onto itself, so loads, the
// pointwise physics before, between and after the transforms, and the stores
// all work on the same registers.  No workgroup barrier runs between the
// table set-up and the FP64 reductions at the end; waves of one workgroup
// proceed independently and overlap their loads with other waves' arithmetic.
#pragma once

#include <type_traits>

#include "gpe_fft.hpp"

namespace gpe {

// lanes per pencil; 0: the length runs in the LDS-tile kernels.  512 = 16
// lanes x 32 points only in complex64 (complex128 would hold 64 registers of
// data per lane and twice the exchange rows)
template <typename T>
constexpr int zw_lanes(int N) {
    switch (N) {
    case 8: return 2;
    case 12: return 2;
    case 16: return 4;
    case 24: return 2;
    case 32: return 4;
    case 48: return 4;
    case 64: return 8;
    case 128: return 8;
    case 256: return 16;
    case 512: return sizeof(T) == 4 ? 16 : 0;
    default: return 0;
    }
}

template <typename T, int N>
struct ZW {
    static constexpr int A = zw_lanes<T>(N);
    static constexpr int B = N / A;
    static constexpr int PW = 64 / A;             // pencils per wave
    static constexpr int ROW = A + 1;             // padded exchange row
    // exchange rows held at once: all B, or A in two phases when B = 2A > 16
    static constexpr int XR = (B > 16) ? A : B;
    // exchange elements per pencil: the smallest odd multiple of A that holds
    // the rows, so the 64/A pencils of a wave start at 32/A different
    // 8-byte bank slots and a wave's exchange access is at most 2-way
    static constexpr int GS = ((XR * ROW + A - 1) / A) % 2 ? ((XR * ROW + A - 1) / A) * A
                                                            : ((XR * ROW + A - 1) / A + 1) * A;
    static constexpr int WAVE = PW * GS;          // exchange elements per wave
    static constexpr int WAVES = kThreads / 64;
    static constexpr int PENCILS = PW * WAVES;    // pencils per workgroup
};

// the largest power-of-two chunk up to cht that divides the B points of a lane
constexpr int zw_chunk(int B, int cht) {
    int c = cht;
    while (c > 1 && B % c != 0) c /= 2;
    return c;
}

// load mode of the energy pass: plain load plus the potential-energy sums
constexpr int LOAD_REDPOT = 3;

// f(integral_constant<int, I>) for I = B .. E-1, each call a separate instance so
// that indices derived from I are compile-time constants
template <int I, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < E) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, E>(f);
    }
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// twiddle w_N^(l kb) of lane l at row kb is entry kb A + l, so the lanes of a
// pencil read consecutive entries

// one transform of sign S on the register layout n = l + A b
template <typename T, int N, int S, int B>
__device__ __forceinline__ void zw_fft(cplx<T> (&v)[B], cplx<T>* X, const TwE<T>* tw, int l) {
    constexpr int A = ZW<T, N>::A, ROW = ZW<T, N>::ROW;
    static_assert(A * B == N, "lane split");
    dft<B, S>(v);
#pragma unroll
    for (int kb = 1; kb < B; kb++) v[kb] = tw_mul<S>(v[kb], tw[kb * A + l]);  // w_N^(l kb)
    if constexpr (B > 16) {
        // two phases through A rows: columns kb < A, then kb >= A.  The second
        // half is in LDS before the first phase's outputs overwrite its registers.
        static_assert(B == 2 * A, "two-phase exchange needs B = 2A");
#pragma unroll
        for (int kb = 0; kb < A; kb++) X[kb * ROW + l] = v[kb];
        wave_lds_sync();
        cplx<T> u[A];
#pragma unroll
        for (int q = 0; q < A; q++) u[q] = X[l * ROW + q];
        wave_lds_sync();
#pragma unroll
        for (int kb = 0; kb < A; kb++) X[kb * ROW + l] = v[A + kb];
        wave_lds_sync();
        dft<A, S>(u);
#pragma unroll
        for (int ka = 0; ka < A; ka++) v[2 * ka] = u[ka];
#pragma unroll
        for (int q = 0; q < A; q++) u[q] = X[l * ROW + q];
        dft<A, S>(u);
#pragma unroll
        for (int ka = 0; ka < A; ka++) v[1 + 2 * ka] = u[ka];
    } else {
#pragma unroll
        for (int kb = 0; kb < B; kb++) X[kb * ROW + l] = v[kb];
        wave_lds_sync();
#pragma unroll
        for (int m = 0; m < B / A; m++) {
            cplx<T> u[A];
#pragma unroll
            for (int q = 0; q < A; q++) u[q] = X[(l + A * m) * ROW + q];
            dft<A, S>(u);
#pragma unroll
            for (int ka = 0; ka < A; ka++) v[m + (B / A) * ka] = u[ka];
        }
    }
    wave_lds_sync();
}

template <typename T, int N>
struct ZwLds {
    TwE<T> tw[N];
    T vz[N];
    cplx<T> x[ZW<T, N>::WAVES * ZW<T, N>::WAVE];
    double red[kThreads / 64];
};

// ---- chunk-vectorised physics for the hot real-time ops.  The branches on
// the physics kind are taken once per chunk, not once per point, and the
// points of a chunk are independent, so the compiler interleaves them.

template <typename T, int CH>
__device__ __forceinline__ void dedfdn_chunk(const Phys<T>& p, const cplx<T> (&v)[CH], T (&out)[CH]) {
    if (p.interactions == INT_CONTACT) {
        const T k = (T)4 * (T)3.14159265358979323846 * p.up[3] * p.gamma;
#pragma unroll
        for (int j = 0; j < CH; j++) out[j] = k * (v[j].x * v[j].x + v[j].y * v[j].y);
    } else if (p.interactions == INT_DIPOLAR) {
#pragma unroll
        for (int j = 0; j < CH; j++) out[j] = (T)0;
    } else {
#pragma unroll
        for (int j = 0; j < CH; j++) out[j] = dedfdn(p, density(p, v[j]), 0u, 0u);
    }
}

// v[j] *= exp(cx lv[j] + i cy lv[j]) (mul_exp per point, the rare large-angle
// case through sincosf for the lanes that need it)
template <int CH>
__device__ __forceinline__ void mul_exp_chunk(cplx<float> (&v)[CH], const float (&lv)[CH], float cx, float cy) {
    float s[CH], c[CH];
    bool big = false;
#pragma unroll
    for (int j = 0; j < CH; j++) {
        const float x = cy * lv[j];
        const float t = x * x;
        s[j] = fmaf(x * t, fmaf(t, fmaf(t, -1.0f / 5040.0f, 1.0f / 120.0f), -1.0f / 6.0f), x);
        c[j] = fmaf(t, fmaf(t, fmaf(t, -1.0f / 720.0f, 1.0f / 24.0f), -0.5f), 1.0f);
        big = big || (fabsf(x) > 0.25f);
    }
    if (big) {
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const float x = cy * lv[j];
            if (fabsf(x) > 0.25f) sincosf(x, &s[j], &c[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < CH; j++) {
        const float h = -0.5f * fmaf(s[j], s[j], fmaf(c[j], c[j], -1.0f));
        const float rx = fmaf(v[j].x, c[j], fmaf(-v[j].y, s[j], h * v[j].x));
        const float ry = fmaf(v[j].y, c[j], fmaf(v[j].x, s[j], h * v[j].y));
        v[j] = mk(rx, ry);
    }
    if (cx != 0.0f) {
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const float gm = expf(cx * lv[j]);
            v[j] = mk(v[j].x * gm, v[j].y * gm);
        }
    }
}
// complex128: the small-angle series for the whole chunk, sincos for the
// lanes past 1/4, exp only when the step has a real part (imaginary or
// complex time)
template <int CH>
__device__ __forceinline__ void mul_exp_chunk(cplx<double> (&v)[CH], const double (&lv)[CH], double cx, double cy) {
    double s[CH], c[CH];
    bool big = false;
#pragma unroll
    for (int j = 0; j < CH; j++) {
        const double x = cy * lv[j];
        sincos_taylor(x, &s[j], &c[j]);
        big = big || (fabs(x) > 0.25);
    }
    if (big) {
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const double x = cy * lv[j];
            if (fabs(x) > 0.25) sincos(x, &s[j], &c[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < CH; j++) v[j] = mk(v[j].x * c[j] - v[j].y * s[j], v[j].x * s[j] + v[j].y * c[j]);
    if (cx != 0.0) {
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const double gm = exp(cx * lv[j]);
            v[j] = mk(v[j].x * gm, v[j].y * gm);
        }
    }
}

// SHAPE: F, I or IF; LM: load op before the first transform; MM: op after it
#ifndef ZW_WAVES_PHYS
#define ZW_WAVES_PHYS 3
#endif
// resident waves per SIMD the compiler must allow (VGPR budget 512 / waves):
// complex64 4, or ZW_WAVES_PHYS for the passes with physics between transforms
// and for the imaginary-time corrector, except the corrector-end pass
// (MID_VSTEP3), which fits 128 VGPRs without spills at every length and runs
// at 4 (256^3: 2780 -> 2818 steps/s; MID_VSTEP2 spills at 4, and with chunks
// of 2 it fits but loses 11 us)
// (512 points: 32 registers of data per lane and three workgroups per CU by
// LDS; 2 waves with physics, 3 without)
template <typename T, int N, int SHAPE, int LM, int MM>
constexpr int zw_waves() {
    constexpr bool phys = (SHAPE == SHAPE_IF && MM != MID_NONE) || LM == LOAD_PART2;
    if constexpr (sizeof(T) != 4) return 2;
    else if constexpr (N > 256) return (phys || MM != MID_NONE || LM != LOAD_PLAIN) ? 2 : 3;
    else return phys ? (MM == MID_VSTEP3 ? 4 : ZW_WAVES_PHYS) : 4;
}

template <typename T, int N, int SHAPE, int LM, int MM>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(zw_waves<T, N, SHAPE, LM, MM>())))
void zwave_kernel(const PassArgs<T> a) {
    using Z = ZW<T, N>;
    constexpr int A = Z::A, B = Z::B;
    __shared__ ZwLds<T, N> L;
    const int vk = a.ph.vext;
    const bool sep = (vk == VEXT_HARMONIC || vk == VEXT_COSINE);
    for (int i = threadIdx.x; i < N; i += kThreads) {
        const int m = (i / A) * (i % A);  // row kb = i / A, lane l = i % A
        if constexpr (sizeof(T) == 4) {
            const cplx<float> h = a.tw[m], lo = a.twlo[m];
            L.tw[i].t = make_float4(h.x, h.y, lo.x, lo.y);
        } else {
            L.tw[i].t = a.tw[m];
        }
        if (sep) L.vz[i] = a.ph.vtab[a.ph.nx + a.ph.ny + i];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane / A, l = lane % A;
    cplx<T>* X = L.x + wv * Z::WAVE + g * Z::GS;
    const unsigned p = (blockIdx.x * Z::WAVES + (unsigned)wv) * Z::PW + (unsigned)g;
    const bool ok = p < a.M;
    const unsigned pp = ok ? p : 0u;
    const unsigned ny = (unsigned)a.ph.ny, nx = (unsigned)a.ph.nx;
    const unsigned bx = pp / ny;
    Coord c;
    c.iy = pp - bx * ny;
    c.ix = bx - (bx / nx) * nx;
    const unsigned base = pp * (unsigned)N + (unsigned)l;
    const T vscale = (vk == VEXT_COSINE) ? a.ph.up[0] : (T)1;
    const T vxy = sep ? a.ph.vtab[c.ix] + a.ph.vtab[a.ph.nx + c.iy] : (T)0;
    const unsigned it = time_index(a.ph);
    double acc[3] = {0.0, 0.0, 0.0};
    auto off_of = [&](int b) __attribute__((always_inline)) { return base + (unsigned)(A * b); };

    // the external potential and the coordinates of register b
    auto elem_at = [&](int b) __attribute__((always_inline)) {
        ElemIn<T> e;
        e.c = c;
        e.it = it;
        e.c.iz = (unsigned)(l + A * b);
        e.c.off = base + (unsigned)(A * b);
        e.vx = sep ? vscale * (vxy + L.vz[l + A * b]) : vext(a.ph, e.c.ix, e.c.iy, e.c.iz, e.c.off, it);
        e.vx1 = e.vx;
        e.aux = mk((T)0, (T)0);
        e.raux = (T)0;
        return e;
    };

    // inactive lanes (past the last pencil) read pencil 0 and store nothing
    cplx<T> v[B];
    {
        const cplx<T>* src = (LM == LOAD_PSI0) ? a.psiA : a.src;
#pragma unroll
        for (int b = 0; b < B; b++) v[b] = src[base + (unsigned)(A * b)];
    }

    // Side inputs (psi0, V1 or V_qf) and the per-point physics run in chunks
    // of CH registers; the loads of chunk q+1 are issued before chunk q is
    // computed, and those of chunk 0 before the first transform, so their
    // latency overlaps arithmetic.
#ifndef ZW_CH
#define ZW_CH 8
#endif
#ifndef ZW_PREFETCH
#define ZW_PREFETCH 1
#endif
    // complex128 holds twice the registers per point, 512-point pencils twice
    // the points per lane: chunks of four
    constexpr int CHT = (sizeof(T) == 4 && N <= 256) ? ZW_CH : 4;
    constexpr int CH = zw_chunk(B, CHT);
    constexpr int NCH = B / CH;
    static_assert(NCH * CH == B, "chunks must tile the points of a lane");
    constexpr bool PRE = (LM != LOAD_PLAIN);   // physics before the first transform
    constexpr bool POST = (MM != MID_NONE);    // physics after it
    constexpr bool NEED_AUX = (LM == LOAD_PART2) || MM == MID_VSTEP2 || MM == MID_PART1 ||
                              MM == MID_QF_VSTEP1 || MM == MID_OVERLAP_RE;
    const T* rs = real_side(a);
    const bool qf = (a.flags & F_QF) != 0;
    // the fast chunk forms: ops without imprint, built-in physics
    const bool fast = (MM == MID_VSTEP2) || (MM == MID_VSTEP3 && !(a.flags & F_IMPRINT_NEXT)) ||
                      (LM == LOAD_PSI0 && !(a.flags & F_IMPRINT)) || (LM == LOAD_PART2 && !(a.flags & F_IMPRINT_MID)) ||
                      MM == MID_PART1 || MM == MID_QF_VSTEP1;
    cplx<T> ax[2][CH];
    T rx[2][CH];
    auto fetch = [&](int q, int slot) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const unsigned off = base + (unsigned)(A * (q * CH + j));
            ax[slot][j] = NEED_AUX ? ld_nt(&a.psiA[off]) : mk((T)0, (T)0);
        }
        if (rs) {
#pragma unroll
            for (int j = 0; j < CH; j++) rx[slot][j] = ld_nt(&rs[off_of(q * CH + j)]);
        } else {
#pragma unroll
            for (int j = 0; j < CH; j++) rx[slot][j] = (T)0;
        }
    };
    auto vx_chunk = [&](int q, T (&vx)[CH]) {
        if (sep) {
#pragma unroll
            for (int j = 0; j < CH; j++) vx[j] = vscale * (vxy + L.vz[l + A * (q * CH + j)]);
        } else {
#pragma unroll
            for (int j = 0; j < CH; j++) {
                const int b = q * CH + j;
                vx[j] = vext(a.ph, c.ix, c.iy, (unsigned)(l + A * b), base + (unsigned)(A * b), it);
            }
        }
    };
    // one chunk of physics on v[q CH .. q CH + CH)
    auto physics = [&](int q, int slot) __attribute__((always_inline)) {
        if (fast) {
            T vx[CH], gv[CH], lv[CH];
            cplx<T> u[CH];
            vx_chunk(q, vx);
#pragma unroll
            for (int j = 0; j < CH; j++) u[j] = v[q * CH + j];
            if constexpr (LM == LOAD_PSI0) {
                // rescale, then Vstep1 (orginal/gpe_engine.cu:723-772)
                if (a.flags & F_SCALE) {
                    const T f = (T)a.scal[1];
#pragma unroll
                    for (int j = 0; j < CH; j++) u[j] = cscale(u[j], f);
                    if (ok) {
#pragma unroll
                        for (int j = 0; j < CH; j++) st_nt(&a.psiA[off_of(q * CH + j)], u[j]);
                    }
                }
                if (a.flags & F_VSTEP1) {
                    dedfdn_chunk(a.ph, u, gv);
#pragma unroll
                    for (int j = 0; j < CH; j++) lv[j] = vx[j] + gv[j] + (qf ? rx[slot][j] : (T)0);
                    mul_exp_chunk(u, lv, a.ph.c.x, a.ph.c.y);
                }
            } else if constexpr (MM == MID_VSTEP2) {
                // predictor end and corrector potential (orginal/gpe_engine.cu:774-801)
                cplx<T> p0[CH];
#pragma unroll
                for (int j = 0; j < CH; j++) p0[j] = ax[slot][j];
                T V0[CH];
                dedfdn_chunk(a.ph, p0, gv);
#pragma unroll
                for (int j = 0; j < CH; j++) V0[j] = vx[j] + gv[j] + (qf ? rx[slot][j] : (T)0);
                mul_exp_chunk(u, V0, a.ph.c.x, a.ph.c.y);  // phi3
                dedfdn_chunk(a.ph, u, gv);
#pragma unroll
                for (int j = 0; j < CH; j++) lv[j] = (T)0.5 * (V0[j] + vx[j] + gv[j] + (qf ? rx[slot][j] : (T)0));
                if (ok) {
#pragma unroll
                    for (int j = 0; j < CH; j++) st_nt(&a.vbuf[off_of(q * CH + j)], lv[j]);
                }
                mul_exp_chunk(p0, lv, a.ph.c.x, a.ph.c.y);
#pragma unroll
                for (int j = 0; j < CH; j++) u[j] = p0[j];
            } else if constexpr (LM == LOAD_PART2) {
                // corrector of imaginary time (orginal/gpe_engine.cu:817-868): u is phi3
                if (a.flags & F_SCALE) {
                    const T f = (T)a.scal[1];
#pragma unroll
                    for (int j = 0; j < CH; j++) u[j] = cscale(u[j], f);
                }
                cplx<T> p0[CH];
                T V0[CH];
#pragma unroll
                for (int j = 0; j < CH; j++) p0[j] = ax[slot][j];
                dedfdn_chunk(a.ph, p0, gv);
#pragma unroll
                for (int j = 0; j < CH; j++) V0[j] = vx[j] + gv[j] + (qf ? rx[slot][j] : (T)0);
                dedfdn_chunk(a.ph, u, gv);
#pragma unroll
                for (int j = 0; j < CH; j++) lv[j] = (T)0.5 * (V0[j] + vx[j] + gv[j] + (qf ? rx[slot][j] : (T)0));
                if (ok) {
#pragma unroll
                    for (int j = 0; j < CH; j++) st_nt(&a.vbuf[off_of(q * CH + j)], lv[j]);
                }
                mul_exp_chunk(p0, lv, a.ph.c.x, a.ph.c.y);
#pragma unroll
                for (int j = 0; j < CH; j++) u[j] = p0[j];
            } else if constexpr (MM == MID_PART1) {
                // predictor end with the density sum (orginal/gpe_engine.cu:803-815)
                cplx<T> p0[CH];
#pragma unroll
                for (int j = 0; j < CH; j++) p0[j] = ax[slot][j];
                dedfdn_chunk(a.ph, p0, gv);
#pragma unroll
                for (int j = 0; j < CH; j++) lv[j] = vx[j] + gv[j] + (qf ? rx[slot][j] : (T)0);
                mul_exp_chunk(u, lv, a.ph.c.x, a.ph.c.y);
                if (ok) {
#pragma unroll
                    for (int j = 0; j < CH; j++) acc[0] += dens64((double)a.ph.gamma, (double)u[j].x, (double)u[j].y);
                }
            } else if constexpr (MM == MID_QF_VSTEP1) {
                // quantum-friction potential from F^-1[qf' k^2 F psi0] (orginal 911-922), then Vstep1_qf
                cplx<T> p0[CH];
#pragma unroll
                for (int j = 0; j < CH; j++) p0[j] = ax[slot][j];
                dedfdn_chunk(a.ph, p0, gv);
#pragma unroll
                for (int j = 0; j < CH; j++) {
                    const T r0 = density(a.ph, p0[j]);
                    const T qv = (p0[j].x * u[j].y + (-p0[j].y) * u[j].x) / (r0 + (T)1.0e-12);
                    if (ok) st_nt(&a.vq[off_of(q * CH + j)], qv);
                    lv[j] = vx[j] + gv[j] + qv;
                }
                mul_exp_chunk(p0, lv, a.ph.c.x, a.ph.c.y);
#pragma unroll
                for (int j = 0; j < CH; j++) u[j] = p0[j];
            } else if constexpr (MM == MID_VSTEP3) {
                // corrector end (orginal/gpe_engine.cu:870-878), then the next step's Vstep1
#pragma unroll
                for (int j = 0; j < CH; j++) lv[j] = rx[slot][j];
                mul_exp_chunk(u, lv, a.ph.c.x, a.ph.c.y);  // psi1
                if ((a.flags & F_RED_NORM) && ok) {
#pragma unroll
                    for (int j = 0; j < CH; j++) acc[0] += dens64((double)a.ph.gamma, (double)u[j].x, (double)u[j].y);
                }
                if (ok) {
#pragma unroll
                    for (int j = 0; j < CH; j++) st_nt(&a.psiA[off_of(q * CH + j)], u[j]);
                }
                if (a.flags & F_NEXT_VSTEP1) {
                    dedfdn_chunk(a.ph, u, gv);
#pragma unroll
                    for (int j = 0; j < CH; j++) lv[j] = vx[j] + gv[j];
                    mul_exp_chunk(u, lv, a.ph.c.x, a.ph.c.y);
                }
            }
#pragma unroll
            for (int j = 0; j < CH; j++) v[q * CH + j] = u[j];
        } else if (ok) {
            constexpr int LMO = (LM == LOAD_REDPOT) ? LOAD_PLAIN : LM;
#pragma unroll
            for (int j = 0; j < CH; j++) {
                ElemIn<T> e = elem_at(q * CH + j);
                e.aux = ax[slot][j];
                e.raux = rx[slot][j];
                if constexpr (PRE) v[q * CH + j] = load_op_m<LMO>(a, e, v[q * CH + j], acc);
                else v[q * CH + j] = mid_op_m<MM>(a, e, v[q * CH + j], acc);
            }
        }
    };

    constexpr int S1 = (SHAPE == SHAPE_F || SHAPE == SHAPE_FI) ? -1 : 1;
    constexpr int S2 = (SHAPE == SHAPE_FI) ? 1 : (SHAPE == SHAPE_IF) ? -1 : 0;
    // complex128: the first chunk's side loads wait for the transform (registers)
    constexpr bool EARLY = ZW_PREFETCH && sizeof(T) == 4;
    if constexpr (PRE || (POST && EARLY)) fetch(0, 0);
    if constexpr (PRE) {
        static_for<0, NCH>([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q + 1 < NCH) fetch(q + 1, (q + 1) & 1);
            physics(q, q & 1);
        });
    }
    zw_fft<T, N, S1, B>(v, X, L.tw, l);
    if constexpr (POST && !EARLY) fetch(0, 0);
    if constexpr (POST) {
        static_for<0, NCH>([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q + 1 < NCH) fetch(q + 1, (q + 1) & 1);
            physics(q, q & 1);
        });
    }
    if constexpr (S2 != 0) zw_fft<T, N, S2, B>(v, X, L.tw, l);

    if (a.store_mode == STORE_PLAIN && ok) {
#pragma unroll
        for (int b = 0; b < B; b++) a.dst[base + (unsigned)(A * b)] = v[b];
    }
    if (a.nred > 0) {
        const double s0 = block_sum(acc[0], L.red);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = s0;
    }
    if (a.nred > 1) {
        const double s1 = block_sum(acc[1], L.red);
        if (threadIdx.x == 0) a.partials[(size_t)gridDim.x + blockIdx.x] = s1;
    }
}

// launches the z pass in the wave kernel when the length and the pass kind
// have an instance; returns the workgroup count, 0 when the caller must use
// the LDS-tile kernels
template <typename T, int N>
unsigned launch_zwave(const PassArgs<T>& a, hipStream_t s) {
    if constexpr (zw_lanes<T>(N) == 0) {
        return 0;
    } else {
        // engines with user hooks run their z passes in the LDS-tile kernels
        // compiled with the hook file; these kernels carry none of its code
        if (a.axis != 2 || a.nred > 2 || a.ph.modify) return 0;
        const unsigned blocks = (a.M + ZW<T, N>::PENCILS - 1) / ZW<T, N>::PENCILS;
        const int shape = pass_shape(a.dir1, a.dir2);
        const int lm = (a.load_mode == LOAD_PLAIN && (a.flags & F_RED_POT)) ? LOAD_REDPOT : a.load_mode;
        const int mm = a.mid_mode;
#define GPE_ZW(SH, LMV, MMV)                                                                          \
    if (shape == SH && lm == LMV && mm == MMV) {                                                      \
        hipLaunchKernelGGL((zwave_kernel<T, N, SH, LMV, MMV>), dim3(blocks), dim3(kThreads), 0, s, a); \
        return blocks;                                                                                \
    }
        GPE_ZW(SHAPE_F, LOAD_PLAIN, MID_NONE)
        GPE_ZW(SHAPE_F, LOAD_REDPOT, MID_NONE)
        GPE_ZW(SHAPE_F, LOAD_PSI0, MID_NONE)
        GPE_ZW(SHAPE_F, LOAD_PART2, MID_NONE)
        GPE_ZW(SHAPE_I, LOAD_PLAIN, MID_NONE)
        GPE_ZW(SHAPE_I, LOAD_PLAIN, MID_PART1)
        GPE_ZW(SHAPE_I, LOAD_PLAIN, MID_VSTEP3)
        GPE_ZW(SHAPE_I, LOAD_PLAIN, MID_OVERLAP_RE)
        GPE_ZW(SHAPE_IF, LOAD_PLAIN, MID_VSTEP2)
        GPE_ZW(SHAPE_IF, LOAD_PLAIN, MID_VSTEP3)
        GPE_ZW(SHAPE_IF, LOAD_PLAIN, MID_QF_VSTEP1)
#undef GPE_ZW
        return 0;
    }
}

}  // namespace gpe
