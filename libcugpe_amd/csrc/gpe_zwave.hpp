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
// again.  Every transform maps this layout onto itself, so loads, the
// pointwise physics before, between and after the transforms, and the stores
// all work on the same registers.  No workgroup barrier runs between the
// table set-up and the FP64 reductions at the end; waves of one workgroup
// proceed independently and overlap their loads with other waves' arithmetic.
#pragma once

#include "gpe_fft.hpp"

namespace gpe {

// lanes per pencil; 0: the length runs in the LDS-tile kernels
constexpr int zw_lanes(int N) {
    switch (N) {
    case 8: return 2;
    case 16: return 4;
    case 32: return 4;
    case 64: return 8;
    case 128: return 8;
    case 256: return 16;
    default: return 0;
    }
}

template <typename T, int N>
struct ZW {
    static constexpr int A = zw_lanes(N);
    static constexpr int B = N / A;
    static constexpr int PW = 64 / A;             // pencils per wave
    static constexpr int ROW = A + 1;             // padded exchange row
    static constexpr int GS = B * ROW + 2;        // exchange elements per pencil
    static constexpr int WAVE = PW * GS;          // exchange elements per wave
    static constexpr int WAVES = kThreads / 64;
    static constexpr int PENCILS = PW * WAVES;    // pencils per workgroup
};

// load mode of the energy pass: plain load plus the potential-energy sums
constexpr int LOAD_REDPOT = 3;
#ifndef ZW_CHUNK
#define ZW_CHUNK 2
#endif

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one transform of sign S on the register layout n = l + A b
template <typename T, int N, int S, int B>
__device__ __forceinline__ void zw_fft(cplx<T> (&v)[B], cplx<T>* X, const cplx<T>* tw, const cplx<T>* twlo,
                                       int l) {
    constexpr int A = ZW<T, N>::A, ROW = ZW<T, N>::ROW;
    static_assert(A * B == N, "lane split");
    dft<B, S>(v);
#pragma unroll
    for (int kb = 1; kb < B; kb++) {  // w_N^(l kb)
        const int m = l * kb;
        cplx<T> t = tw[m];
        if (S > 0) t.y = -t.y;
        if constexpr (sizeof(T) == 4) {
            cplx<T> tl = twlo[m];
            if (S > 0) tl.y = -tl.y;
            v[kb] = cmul_comp(v[kb], t, tl);
        } else {
            v[kb] = cmul(v[kb], t);
        }
    }
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
    wave_lds_sync();
}

template <typename T, int N>
struct ZwLds {
    cplx<T> tw[N];
    cplx<T> twlo[sizeof(T) == 4 ? N : 1];
    T vz[N];
    cplx<T> x[ZW<T, N>::WAVES * ZW<T, N>::WAVE];
    double red[kThreads / 64];
};

// SHAPE: F, I or IF; LM: load op before the first transform; MM: op after it
template <typename T, int N, int SHAPE, int LM, int MM>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
void zwave_kernel(const PassArgs<T> a) {
    using Z = ZW<T, N>;
    constexpr int A = Z::A, B = Z::B;
    __shared__ ZwLds<T, N> L;
    const int vk = a.ph.vext;
    const bool sep = (vk == VEXT_HARMONIC || vk == VEXT_COSINE);
    for (int i = threadIdx.x; i < N; i += kThreads) {
        L.tw[i] = a.tw[i];
        if constexpr (sizeof(T) == 4) L.twlo[i] = a.twlo[i];
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
    double acc[3] = {0.0, 0.0, 0.0};

    // the external potential and the coordinates of register b
    auto elem_at = [&](int b) {
        ElemIn<T> e;
        e.c = c;
        e.c.iz = (unsigned)(l + A * b);
        e.c.off = base + (unsigned)(A * b);
        e.vx = sep ? vscale * (vxy + L.vz[l + A * b]) : vext(a.ph, e.c.ix, e.c.iy, e.c.iz, e.c.off);
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
    if constexpr (LM != LOAD_PLAIN) {
        // LOAD_REDPOT: the energy sums of rho V_ext and EDF(rho) on the loaded psi
        constexpr int LMO = (LM == LOAD_REDPOT) ? LOAD_PLAIN : LM;
        const T* rs = real_side(a);
#pragma unroll
        for (int b0 = 0; b0 < B; b0 += ZW_CHUNK) {
            cplx<T> aux[ZW_CHUNK];
            T raux[ZW_CHUNK];
#pragma unroll
            for (int q = 0; q < ZW_CHUNK; q++) {
                const unsigned off = base + (unsigned)(A * (b0 + q));
                aux[q] = (LM == LOAD_PART2) ? a.psiA[off] : mk((T)0, (T)0);
                raux[q] = rs ? rs[off] : (T)0;
            }
            if (ok) {
#pragma unroll
                for (int q = 0; q < ZW_CHUNK; q++) {
                    ElemIn<T> e = elem_at(b0 + q);
                    e.aux = aux[q];
                    e.raux = raux[q];
                    v[b0 + q] = load_op_m<LMO>(a, e, v[b0 + q], acc);
                }
            }
        }
    }

    constexpr int S1 = (SHAPE == SHAPE_F || SHAPE == SHAPE_FI) ? -1 : 1;
    constexpr int S2 = (SHAPE == SHAPE_FI) ? 1 : (SHAPE == SHAPE_IF) ? -1 : 0;
    zw_fft<T, N, S1, B>(v, X, L.tw, L.twlo, l);

    if constexpr (MM != MID_NONE) {
        constexpr bool want_psi0 = (MM == MID_VSTEP2 || MM == MID_PART1 || MM == MID_QF_VSTEP1 ||
                                    MM == MID_OVERLAP_RE);
        const T* rs = real_side(a);
#pragma unroll
        for (int b0 = 0; b0 < B; b0 += ZW_CHUNK) {
            cplx<T> aux[ZW_CHUNK];
            T raux[ZW_CHUNK];
#pragma unroll
            for (int q = 0; q < ZW_CHUNK; q++) {
                const unsigned off = base + (unsigned)(A * (b0 + q));
                aux[q] = want_psi0 ? a.psiA[off] : mk((T)0, (T)0);
                raux[q] = rs ? rs[off] : (T)0;
            }
            if (ok) {
#pragma unroll
                for (int q = 0; q < ZW_CHUNK; q++) {
                    ElemIn<T> e = elem_at(b0 + q);
                    e.aux = aux[q];
                    e.raux = raux[q];
                    v[b0 + q] = mid_op_m<MM>(a, e, v[b0 + q], acc);
                }
            }
        }
    }
    if constexpr (S2 != 0) zw_fft<T, N, S2, B>(v, X, L.tw, L.twlo, l);

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
    if constexpr (zw_lanes(N) == 0) {
        return 0;
    } else {
        if (a.axis != 2 || a.nred > 2) return 0;
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
