// Whole real-time split steps of small lattices in one kernel launch.
//
// A lattice of L^3 points whose state fits in one workgroup's LDS (complex64
// up to 24^3 = 110.6 KB, complex128 up to 16^3 = 64 KB) is advanced by one
// workgroup for all nt steps of a launch; HBM is touched only to load psi at
// the start and store it at the end.  The step is the one the axis passes run
// (orginal/gpe_engine.cu:942-1016, real time, no quantum friction):
//
//   phase A  psi0 -> Vstep1 -> forward z         (pair of lanes = z pencil)
//   phase Y  forward y, K_y K_z, inverse y        (pair of lanes = y pencil)
//   phase X  forward x, K_x, inverse x            (pair of lanes = x pencil)
//   phase B  inverse z -> Vstep2 -> forward z     (keeps V1 in registers)
//   phase Y, phase X
//   phase C  inverse z -> Vstep3 = psi1 -> next step's phase A
//
// Two adjacent lanes own each pencil, L/2 points each, and transform it in
// registers: an L/2-point DFT per lane plus one radix-2 step across the pair
// (a lane shuffle).  A forward decimation in time takes the interleaved
// layout (lane h holds points 2j + h) to the halves layout (lane h holds
// L/2 h + j); the decimation in frequency takes halves back to interleaved.
// psi0 and V1 of the physics phases stay in registers, interleaved, across
// the kinetic phases.  The LDS cube is [ix][iy][iz] with z rows padded to
// L + 1 elements, so that the z pencils of a wave start on different banks
// and the y and x pencils of a wave are nearly consecutive.
#pragma once

#include "gpe_zwave.hpp"

namespace gpe {

template <typename T>
struct FusedArgs {
    cplx<T>* psi;                // batch lattices of L^3 points, one after another
    int nt;                      // steps of this launch
    unsigned it0;                // time index of the first step
    const cplx<double>* Kx;      // FP64 kinetic factors (Kz includes 1/N)
    const cplx<double>* Ky;
    const cplx<double>* Kz;
    const cplx<float>* kx_hl;    // complex64: K_x as (hi, lo) float pairs
    const cplx<float>* kyz_hl;   // complex64: K_y K_z per (iy, iz) as (hi, lo) pairs
    Phys<T> ph;
};

// z pencils per lane pair: two for the complex64 24^3 cube, so that its
// workgroup of 576 threads keeps psi0 and V1 of two pencils per pair in
// registers at three waves per SIMD; one otherwise
template <typename T, int L>
constexpr int fused_pp() {
    return (sizeof(T) == 4 && L == 24) ? 2 : 1;
}

// threads of a workgroup: a lane pair per group of fused_pp z pencils,
// rounded up to whole waves
template <typename T, int L>
constexpr int fused_threads() {
    return ((2 * L * L / fused_pp<T, L>() + 63) / 64) * 64;
}

template <typename T, int L>
struct FusedLds {
    static constexpr int PY = L + 1;      // z rows padded by one element
    static constexpr int PX = L * PY;
    static constexpr bool C64 = sizeof(T) == 4;
    cplx<T> x[L * PX];
    float4 kyz[C64 ? L * L : 1];          // complex64: K_y K_z hi.x, hi.y, lo.x, lo.y
    float4 kx[C64 ? L : 1];
    cplx<double> K[C64 ? 1 : 3 * L];      // complex128: K_x | K_y | K_z
    T vt[3 * L];                          // per-axis terms of a separable V_ext
};

// physics chunk: the largest of 6, 4 dividing L / 2
template <int M>
constexpr int fused_chunk() {
    return (M % 6 == 0) ? 6 : (M % 4 == 0) ? 4 : 2;
}

// v[j] *= exp(S 2 pi i j / L) on the odd lane of a pair
template <int L, int S, int J, int M, typename T>
__device__ __forceinline__ void fused_twiddle(cplx<T> (&v)[M], bool odd) {
    if constexpr (J < M) {
        if (odd) v[J] = rot<L, J, S>(v[J]);
        fused_twiddle<L, S, J + 1, M>(v, odd);
    }
}

template <typename T>
__device__ __forceinline__ cplx<T> pair_swap(cplx<T> v) {
    return mk(__shfl_xor(v.x, 1), __shfl_xor(v.y, 1));
}

// L-point DFT (sign S) of a pencil held by a lane pair, interleaved -> halves
template <int L, int S, typename T>
__device__ __forceinline__ void fused_dit(cplx<T> (&v)[L / 2], bool odd) {
    constexpr int M = L / 2;
    dft<M, S>(v);
    fused_twiddle<L, S, 0, M>(v, odd);
#pragma unroll
    for (int k = 0; k < M; k++) {
        const cplx<T> o = pair_swap(v[k]);
        v[k] = odd ? csub(o, v[k]) : cadd(v[k], o);
    }
}

// L-point DFT (sign S) of a pencil held by a lane pair, halves -> interleaved
template <int L, int S, typename T>
__device__ __forceinline__ void fused_dif(cplx<T> (&v)[L / 2], bool odd) {
    constexpr int M = L / 2;
#pragma unroll
    for (int j = 0; j < M; j++) {
        const cplx<T> o = pair_swap(v[j]);
        v[j] = odd ? csub(o, v[j]) : cadd(v[j], o);
    }
    fused_twiddle<L, S, 0, M>(v, odd);
    dft<M, S>(v);
}

// V_ext at (ix, iy, iz) from the LDS per-axis terms when the potential is
// separable
template <bool UH, typename T, int L>
__device__ __forceinline__ T fused_vext(const Phys<T>& p, const FusedLds<T, L>& S, bool sep, T vscale, unsigned ix,
                                        unsigned iy, unsigned iz, unsigned off, unsigned it) {
    if (!UH && sep) return vscale * (S.vt[ix] + S.vt[L + iy] + S.vt[2 * L + iz]);
    return vext<UH>(p, ix, iy, iz, off, it);
}

// the y (AXIS 1) or x (AXIS 0) turn-around of the pencil at `base` with
// element stride `st`: forward transform, kinetic factor, inverse transform
template <typename T, int L, int AXIS>
__device__ __forceinline__ void fused_turn(FusedLds<T, L>& S, int base, int st, int col, int h) {
    constexpr int M = L / 2;
    const bool odd = h != 0;
    cplx<T> u[M];
#pragma unroll
    for (int j = 0; j < M; j++) u[j] = S.x[base + (2 * j + h) * st];
    fused_dit<L, -1>(u, odd);
    // keep the factor-table reads here: hoisted above the transform they would
    // hold registers through it (a compiler-only barrier, no instruction)
    asm volatile("" ::: "memory");
#pragma unroll
    for (int j = 0; j < M; j++) {
        const int k = M * h + j;
        if constexpr (sizeof(T) == 4) {
            // one compensated product with the hi + lo pair of an FP64 factor
            const float4 q = (AXIS == 1) ? S.kyz[k * L + col] : S.kx[k];
            u[j] = cmul_comp(u[j], mk(q.x, q.y), mk(q.z, q.w));
        } else {
            if constexpr (AXIS == 1) u[j] = cmul(cmul(u[j], S.K[L + k]), S.K[2 * L + col]);
            else u[j] = cmul(u[j], S.K[k]);
        }
    }
    fused_dif<L, +1>(u, odd);
#pragma unroll
    for (int j = 0; j < M; j++) S.x[base + (2 * j + h) * st] = u[j];
}

template <typename T, int L, bool UH>
__global__ __launch_bounds__((fused_threads<T, L>())) void fused_steps_kernel(const FusedArgs<T> a) {
    using FL = FusedLds<T, L>;
    __shared__ FL S;
    constexpr int NP = L * L;
    constexpr int PP = fused_pp<T, L>();
    constexpr int NT = fused_threads<T, L>();
    constexpr int M = L / 2;
    constexpr int CH = fused_chunk<M>();
    const int t = threadIdx.x;
    const Phys<T>& ph = a.ph;
    const int vk = ph.vext;
    const bool sep = (vk == VEXT_HARMONIC || vk == VEXT_COSINE);
    const T vscale = (vk == VEXT_COSINE) ? ph.up[0] : (T)1;
    for (int i = t; i < NP; i += NT) {
        if constexpr (FL::C64) {
            const cplx<float> hi = a.kyz_hl[2 * i], lo = a.kyz_hl[2 * i + 1];
            S.kyz[i] = make_float4(hi.x, hi.y, lo.x, lo.y);
        }
    }
    for (int i = t; i < 3 * L; i += NT) {
        if constexpr (FL::C64) {
            if (i < L) {
                const cplx<float> hi = a.kx_hl[2 * i], lo = a.kx_hl[2 * i + 1];
                S.kx[i] = make_float4(hi.x, hi.y, lo.x, lo.y);
            }
        } else {
            S.K[i] = (i < L) ? a.Kx[i] : (i < 2 * L) ? a.Ky[i - L] : a.Kz[i - 2 * L];
        }
        S.vt[i] = sep ? ph.vtab[i] : (T)0;
    }

    // Lane pair p owns pencils p + c NP / PP (c < PP) of every axis: z
    // (ix, iy) for the physics, y (ix, iz) and x (iy, iz) for the kinetic
    // factor; both lanes of a pair agree on `own`.
    const int p = t >> 1, h = t & 1;
    const bool odd = h != 0;
    const bool own = p < NP / PP;
    int r0[PP], r1[PP];
#pragma unroll
    for (int c = 0; c < PP; c++) {
        const int q = own ? p + c * (NP / PP) : 0;
        r0[c] = q / L;
        r1[c] = q % L;
    }
    const unsigned n1 = (unsigned)(L * L * L);
    auto gbase = [&](int c) { return blockIdx.x * n1 + (unsigned)(r0[c] * L + r1[c]) * L; };
    auto zrow = [&](int c) { return r0[c] * FL::PX + r1[c] * FL::PY; };
    const T cx = ph.c.x, cy = ph.c.y;

    cplx<T> p0[PP][M];  // psi0 of the step at z = 2 j + h
    T v1[PP][M];        // corrector potential V1 there
#pragma unroll
    for (int c = 0; c < PP; c++) {
#pragma unroll
        for (int j = 0; j < M; j++) p0[c][j] = own ? a.psi[gbase(c) + 2 * j + h] : mk((T)0, (T)0);
    }
    __syncthreads();

    // the y and x turn-arounds of all the pair's pencils, each after a barrier
    auto kinetic = [&]() __attribute__((always_inline)) {
        __syncthreads();
        if (own) {
#pragma unroll
            for (int c = 0; c < PP; c++) fused_turn<T, L, 1>(S, r0[c] * FL::PX + r1[c], FL::PY, r1[c], h);
        }
        __syncthreads();
        if (own) {
#pragma unroll
            for (int c = 0; c < PP; c++) fused_turn<T, L, 0>(S, r0[c] * FL::PY + r1[c], FL::PX, 0, h);
        }
        __syncthreads();
    };

    unsigned it = a.it0;
    for (int s = 0; s < a.nt; s++, it++) {
        // phase A: Vstep1 (orginal/gpe_engine.cu:723-772), forward z
#pragma unroll
        for (int c = 0; c < PP; c++) {
            cplx<T> u[M];
            const unsigned gb = gbase(c);
            if constexpr (!UH) {
#pragma unroll
                for (int q = 0; q < M / CH; q++) {
                    cplx<T> w[CH];
                    T gv[CH], lv[CH];
#pragma unroll
                    for (int j = 0; j < CH; j++) w[j] = p0[c][q * CH + j];
                    dedfdn_chunk(ph, w, gv);
#pragma unroll
                    for (int j = 0; j < CH; j++) {
                        const unsigned iz = (unsigned)(2 * (q * CH + j) + h);
                        lv[j] = fused_vext<UH>(ph, S, sep, vscale, r0[c], r1[c], iz, gb + iz, it) + gv[j];
                    }
                    mul_exp_chunk(w, lv, cx, cy);
#pragma unroll
                    for (int j = 0; j < CH; j++) u[q * CH + j] = w[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < M; j++) {
                    const unsigned iz = (unsigned)(2 * j + h), off = gb + iz;
                    const cplx<T> m0 = modify_psi<UH>(ph, p0[c][j], r0[c], r1[c], iz, off, it);
                    const T lv = fused_vext<UH>(ph, S, sep, vscale, r0[c], r1[c], iz, off, it) +
                                 dedfdn<UH>(ph, density(ph, m0), it, off);
                    u[j] = mul_exp(m0, cx * lv, cy * lv);
                }
            }
            fused_dit<L, -1>(u, odd);
            if (own) {
#pragma unroll
                for (int j = 0; j < M; j++) S.x[zrow(c) + M * h + j] = u[j];
            }
        }
        kinetic();

        // phase B: inverse z, Vstep2 (orginal/gpe_engine.cu:774-801), forward z
#pragma unroll
        for (int c = 0; c < PP; c++) {
            cplx<T> u[M];
            const unsigned gb = gbase(c);
#pragma unroll
            for (int j = 0; j < M; j++) u[j] = S.x[zrow(c) + M * h + j];
            fused_dif<L, +1>(u, odd);
            if constexpr (!UH) {
#pragma unroll
                for (int q = 0; q < M / CH; q++) {
                    cplx<T> w[CH], z0[CH];
                    T gv[CH], V0[CH], vx[CH];
#pragma unroll
                    for (int j = 0; j < CH; j++) {
                        const unsigned iz = (unsigned)(2 * (q * CH + j) + h);
                        w[j] = u[q * CH + j];
                        z0[j] = p0[c][q * CH + j];
                        vx[j] = fused_vext<UH>(ph, S, sep, vscale, r0[c], r1[c], iz, gb + iz, it);
                    }
                    dedfdn_chunk(ph, z0, gv);
#pragma unroll
                    for (int j = 0; j < CH; j++) V0[j] = vx[j] + gv[j];
                    mul_exp_chunk(w, V0, cx, cy);  // phi3
                    dedfdn_chunk(ph, w, gv);
#pragma unroll
                    for (int j = 0; j < CH; j++) {
                        const T l = (T)0.5 * (V0[j] + vx[j] + gv[j]);
                        v1[c][q * CH + j] = l;
                        V0[j] = l;
                    }
                    mul_exp_chunk(z0, V0, cx, cy);
#pragma unroll
                    for (int j = 0; j < CH; j++) u[q * CH + j] = z0[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < M; j++) {
                    const unsigned iz = (unsigned)(2 * j + h), off = gb + iz;
                    const cplx<T> m0 = modify_psi<UH>(ph, p0[c][j], r0[c], r1[c], iz, off, it);
                    const T V0 = fused_vext<UH>(ph, S, sep, vscale, r0[c], r1[c], iz, off, it) +
                                 dedfdn<UH>(ph, density(ph, m0), it, off);
                    const cplx<T> phi3 = mul_exp(u[j], cx * V0, cy * V0);
                    const T l = (T)0.5 * (V0 + fused_vext<UH>(ph, S, sep, vscale, r0[c], r1[c], iz, off, it + 1) +
                                          dedfdn<UH>(ph, density(ph, phi3), it + 1, off));
                    v1[c][j] = l;
                    u[j] = mul_exp(m0, cx * l, cy * l);
                }
            }
            fused_dit<L, -1>(u, odd);
            if (own) {
#pragma unroll
                for (int j = 0; j < M; j++) S.x[zrow(c) + M * h + j] = u[j];
            }
        }
        kinetic();

        // phase C: inverse z, Vstep3 (orginal/gpe_engine.cu:870-878) gives psi1
#pragma unroll
        for (int c = 0; c < PP; c++) {
            cplx<T> u[M];
#pragma unroll
            for (int j = 0; j < M; j++) u[j] = S.x[zrow(c) + M * h + j];
            fused_dif<L, +1>(u, odd);
            if constexpr (!UH) {
#pragma unroll
                for (int q = 0; q < M / CH; q++) {
                    cplx<T> w[CH];
                    T lv[CH];
#pragma unroll
                    for (int j = 0; j < CH; j++) {
                        w[j] = u[q * CH + j];
                        lv[j] = v1[c][q * CH + j];
                    }
                    mul_exp_chunk(w, lv, cx, cy);
#pragma unroll
                    for (int j = 0; j < CH; j++) p0[c][q * CH + j] = w[j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < M; j++) p0[c][j] = mul_exp(u[j], cx * v1[c][j], cy * v1[c][j]);
            }
        }
        // the next phase A writes only the pair's own z rows, which no other
        // pair reads before the barrier that follows it
    }
    if (own) {
#pragma unroll
        for (int c = 0; c < PP; c++) {
#pragma unroll
            for (int j = 0; j < M; j++) a.psi[gbase(c) + 2 * j + h] = p0[c][j];
        }
    }
}

}  // namespace gpe
