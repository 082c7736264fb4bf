// y-z plane passes: the z and y transforms of a whole plane (fixed lattice
// and ix) in one workgroup's LDS, with the step's physics in between.
//
// When a plane of ny = nz = L points fits in LDS (L = 32, 48, 64 in both
// precisions), a real-time step (orginal/gpe_engine.cu:942-1016) needs four
// HBM passes instead of six, the kinetic factor K_x K_y K_z / N split as in
// the axis passes:
//
//   open       psi0 -> Vstep1 -> forward z, forward y, K_y K_z     (A -> B)
//   step       x turn-around on B (forward x, K_x, inverse x)
//              inverse y, inverse z, Vstep2 (psi0 from A, V1 out),
//                forward z, forward y, K_y K_z                     (B -> A)
//              x turn-around on A
//              inverse y, inverse z, Vstep3 (psi1 to A) and the
//                next step's Vstep1, forward z, forward y, K_y K_z (A -> B)
//              (the last step ends after Vstep3)
//
// 11 S of traffic per step instead of 15 S.  The plane is stored [iy][iz]
// with rows padded to L + 1 elements; a Stockham stage takes pencil w and
// butterfly j from b = j * L + w, so that the lanes of a wave walk pencils:
// rows of stride L + 1 for the z transforms, consecutive elements for the y
// transforms, both free of bank conflicts.  The physics is the per-element
// code of the LDS-tile passes (load_op_m / mid_op_m).
#pragma once

#include "gpe_fft.hpp"

namespace gpe {

enum PlaneMode { PL_OPEN = 0, PL_VSTEP2 = 1, PL_VSTEP3_NEXT = 2, PL_VSTEP3_LAST = 3 };

constexpr int kPlaneThreads = 256;

template <typename T, int L>
struct PlaneLds {
    static constexpr int P = L + 1;
    cplx<T> x[L * P];
    TwE<T> tw[L];
};

// radix-R Stockham stage (span NS) of L pencils of L points; element i of
// pencil w sits at w * WS + i * IS
template <typename T, int L, int WS, int IS, int S, int R, int NS>
__device__ __forceinline__ void plane_stage(cplx<T>* x, const TwE<T>* tw) {
    constexpr int M = L / R;
    constexpr int NB = L * M;
    constexpr int PER = (NB + kPlaneThreads - 1) / kPlaneThreads;
    constexpr int TSTEP = L / (NS * R);
    cplx<T> v[PER][R];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kPlaneThreads;
        if ((NB % kPlaneThreads) == 0 || b < NB) {
            const int w = b % L, j = b / L;
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = x[w * WS + (j + r * M) * IS];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kPlaneThreads;
        if ((NB % kPlaneThreads) == 0 || b < NB) {
            const int w = b % L, j = b / L;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[q][r] = tw_mul<S>(v[q][r], tw[k * r * TSTEP]);
            }
            dft<R, S>(v[q]);
            const int d = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) x[w * WS + (d + r * NS) * IS] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int WS, int IS, int S, int STAGE>
__device__ __forceinline__ void plane_transform(cplx<T>* x, const TwE<T>* tw) {
    constexpr FftPlan P = plan_of(L, false, sizeof(T) == 4);
    if constexpr (STAGE < P.ns) {
        plane_stage<T, L, WS, IS, S, P.r[STAGE], plan_span(P, STAGE)>(x, tw);
        plane_transform<T, L, WS, IS, S, STAGE + 1>(x, tw);
    }
}

// 2-D transform of sign S over the plane: z (along rows), then y
template <typename T, int L, int S>
__device__ __forceinline__ void plane_fft2(cplx<T>* x, const TwE<T>* tw) {
    constexpr int P = L + 1;
    plane_transform<T, L, P, 1, S, 0>(x, tw);
    plane_transform<T, L, 1, P, S, 0>(x, tw);
}

// inputs of the physics at element (iy, iz) of plane (b, ix)
template <bool UH, typename T>
__device__ __forceinline__ ElemIn<T> plane_elem(const PassArgs<T>& a, unsigned ix, unsigned iy, unsigned iz,
                                                unsigned off, unsigned it, bool sep, T vscale, bool psi0, bool v1) {
    ElemIn<T> e;
    e.it = it;
    e.c.ix = ix;
    e.c.iy = iy;
    e.c.iz = iz;
    e.c.off = off;
    const int nx = a.ph.nx, ny = a.ph.ny;
    e.vx = sep ? vscale * (a.ph.vtab[ix] + a.ph.vtab[nx + iy] + a.ph.vtab[nx + ny + iz])
               : vext<UH>(a.ph, ix, iy, iz, off, it);
    e.vx1 = e.vx;
    if constexpr (UH) {
        if (a.ph.vext == VEXT_USER) e.vx1 = vext<UH>(a.ph, ix, iy, iz, off, it + 1);
    }
    e.aux = psi0 ? a.psiA[off] : mk((T)0, (T)0);
    e.raux = v1 ? a.vbuf[off] : (T)0;
    return e;
}

// One plane pass (MODE: PlaneMode) over every plane of the batch: workgroup
// blockIdx.x takes plane (b, ix) = (blockIdx.x / nx, blockIdx.x % nx).
template <typename T, int L, int MODE, bool UH>
__global__ __launch_bounds__(kPlaneThreads) __attribute__((amdgpu_waves_per_eu(2)))
void plane_kernel(const PassArgs<T> a) {
    constexpr int P = L + 1;
    constexpr int NPT = L * L;
    __shared__ PlaneLds<T, L> S;
    for (int i = threadIdx.x; i < L; i += kPlaneThreads) {
        if constexpr (sizeof(T) == 4) {
            const cplx<float> h = a.tw[i], lo = a.twlo[i];
            S.tw[i].t = make_float4(h.x, h.y, lo.x, lo.y);
        } else {
            S.tw[i].t = a.tw[i];
        }
    }
    const unsigned nx = (unsigned)a.ph.nx;
    const unsigned pl = blockIdx.x;
    const unsigned ix = pl % nx;
    const unsigned base = (pl / nx) * a.n1 + ix * (unsigned)NPT;
    const unsigned it = time_index(a.ph);
    const int vk = a.ph.vext;
    const bool sep = (vk == VEXT_HARMONIC || vk == VEXT_COSINE);
    const T vscale = (vk == VEXT_COSINE) ? a.ph.up[0] : (T)1;
    double acc[3] = {0.0, 0.0, 0.0};

    if constexpr (MODE == PL_OPEN) {
        // psi0 with the step's Vstep1 (the load op of the open z pass)
        for (int e = threadIdx.x; e < NPT; e += kPlaneThreads) {
            const unsigned iy = (unsigned)(e / L), iz = (unsigned)(e % L), off = base + (unsigned)e;
            const ElemIn<T> in = plane_elem<UH>(a, ix, iy, iz, off, it, sep, vscale, false, false);
            S.x[iy * P + iz] = load_op_m<LOAD_PSI0, UH>(a, in, a.psiA[off], acc);
        }
        __syncthreads();
    } else {
        for (int e = threadIdx.x; e < NPT; e += kPlaneThreads) S.x[(e / L) * P + e % L] = a.src[base + (unsigned)e];
        __syncthreads();
        plane_fft2<T, L, +1>(S.x, S.tw);
        for (int e = threadIdx.x; e < NPT; e += kPlaneThreads) {
            const unsigned iy = (unsigned)(e / L), iz = (unsigned)(e % L), off = base + (unsigned)e;
            cplx<T>& v = S.x[iy * P + iz];
            if constexpr (MODE == PL_VSTEP2) {
                v = mid_op_m<MID_VSTEP2, UH>(a, plane_elem<UH>(a, ix, iy, iz, off, it, sep, vscale, true, false), v,
                                             acc);
            } else {
                v = mid_op_m<MID_VSTEP3, UH>(a, plane_elem<UH>(a, ix, iy, iz, off, it, sep, vscale, false, true), v,
                                             acc);
            }
        }
        if constexpr (MODE == PL_VSTEP3_LAST) return;  // psi1 is in psi (A)
        __syncthreads();
    }
    plane_fft2<T, L, -1>(S.x, S.tw);
    // K_y K_z / N of the split kinetic factor, as the y turn-around pass applies it
    for (int e = threadIdx.x; e < NPT; e += kPlaneThreads) {
        Coord c;
        c.iy = (unsigned)(e / L);
        c.iz = (unsigned)(e % L);
        c.ix = ix;
        c.off = base + (unsigned)e;
        a.dst[c.off] = mid_op_lean_k<MK_SEP, false, T, 1>(a, c, S.x[c.iy * P + c.iz]);
    }
}

}  // namespace gpe
