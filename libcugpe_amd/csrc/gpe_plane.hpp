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
// transforms, both free of bank conflicts.  The physics runs in registers on
// the outputs of the inverse y transform's last stage, in the chunked form of
// the z-wave passes (per element with the user hooks), and the kinetic factor
// K_y K_z / N and the store on those of the forward y transform's last stage,
// the factors K_y and K_z / N taken from LDS tables (complex64: two
// compensated products with float hi + lo pairs).
#pragma once

#include "gpe_zwave.hpp"

namespace gpe {

enum PlaneMode { PL_OPEN = 0, PL_VSTEP2 = 1, PL_VSTEP3_NEXT = 2, PL_VSTEP3_LAST = 3 };

// threads of a plane workgroup: 512 for 64 x 64 planes (one butterfly of the
// last stage per thread), 256 otherwise
template <int L>
constexpr int plane_threads() {
    return L >= 64 ? 512 : 256;
}

template <typename T, int L>
struct PlaneLds {
    static constexpr int P = L + 1;
    cplx<T> x[L * P];
    TwE<T> tw[L];
    TwE<T> ky[L];  // K_y and K_z / N (complex64: float hi + lo pairs of the FP64 factors)
    TwE<T> kz[L];
};

// element (iy, iz) times K_y[iy] K_z[iz] / N from the LDS tables: complex64
// as two compensated products, complex128 as the FP64 products of the y
// turn-around pass
template <typename T, int L>
__device__ __forceinline__ cplx<T> plane_kin(const PlaneLds<T, L>& S, unsigned iy, unsigned iz, cplx<T> v) {
    if constexpr (sizeof(T) == 4) {
        const float4 y = S.ky[iy].t, z = S.kz[iz].t;
        v = cmul_comp(v, mk(y.x, y.y), mk(y.z, y.w));
        return cmul_comp(v, mk(z.x, z.y), mk(z.z, z.w));
    } else {
        return cmul(cmul(v, S.ky[iy].t), S.kz[iz].t);
    }
}

// radix-R Stockham stage (span NS) of L pencils of L points; element i of
// pencil w sits at w * WS + i * IS
template <typename T, int L, int WS, int IS, int S, int R, int NS>
__device__ __forceinline__ void plane_stage(cplx<T>* x, const TwE<T>* tw) {
    constexpr int kPlaneThreads = plane_threads<L>();
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

// stages [STAGE, LAST) of the plan of length L
template <typename T, int L, int WS, int IS, int S, int STAGE, int LAST>
__device__ __forceinline__ void plane_stages(cplx<T>* x, const TwE<T>* tw) {
    constexpr FftPlan P = plan_of(L, false, sizeof(T) == 4);
    if constexpr (STAGE < LAST) {
        plane_stage<T, L, WS, IS, S, P.r[STAGE], plan_span(P, STAGE)>(x, tw);
        plane_stages<T, L, WS, IS, S, STAGE + 1, LAST>(x, tw);
    }
}

// radix of the last stage and the butterflies per thread there
template <typename T, int L>
struct PlaneLast {
    static constexpr int kPlaneThreads = plane_threads<L>();
    static constexpr FftPlan P = plan_of(L, false, sizeof(T) == 4);
    static constexpr int R = P.r[P.ns - 1];
    static constexpr int M = L / R;
    static constexpr int NB = L * M;
    static constexpr int PER = (NB + kPlaneThreads - 1) / kPlaneThreads;
};

// Last stage of a y transform (element iy of pencil iz at iz + iy (L + 1))
// into registers: v[q][r] is element iy = j + r M of pencil iz = w, with
// (w, j) of butterfly b = threadIdx.x + q * kPlaneThreads; ok[q] marks the
// butterflies that exist.  Every thread's LDS reads are done on return.
template <typename T, int L, int S>
__device__ __forceinline__ void plane_last_y(const cplx<T>* x, const TwE<T>* tw,
                                             cplx<T> (&v)[PlaneLast<T, L>::PER][PlaneLast<T, L>::R]) {
    using PL = PlaneLast<T, L>;
    constexpr int kPlaneThreads = plane_threads<L>();
    constexpr int R = PL::R, M = PL::M, P = L + 1;
#pragma unroll
    for (int q = 0; q < PL::PER; q++) {
        const int b = threadIdx.x + q * kPlaneThreads;
        if ((PL::NB % kPlaneThreads) == 0 || b < PL::NB) {
            const int w = b % L, j = b / L;
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = x[w + (j + r * M) * P];
#pragma unroll
            for (int r = 1; r < R; r++) v[q][r] = tw_mul<S>(v[q][r], tw[j * r]);
            dft<R, S>(v[q]);
        }
    }
    __syncthreads();
}

// inputs of the per-element physics at element (iy, iz) of plane (b, ix)
// (the user hook kernels)
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

// chunked V_ext over the R elements (iy = j + r M, iz = w) of a butterfly
template <bool UH, typename T, int R, int M>
__device__ __forceinline__ void plane_vext(const PassArgs<T>& a, unsigned ix, int w, int j, unsigned base, unsigned it,
                                           bool sep, T vscale, T (&vx)[R]) {
    const int nx = a.ph.nx, ny = a.ph.ny;
    if (!UH && sep) {
        const T vxz = a.ph.vtab[ix];
        const T vz = a.ph.vtab[nx + ny + w];
#pragma unroll
        for (int r = 0; r < R; r++) vx[r] = vscale * (vxz + a.ph.vtab[nx + j + r * M] + vz);
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const unsigned iy = (unsigned)(j + r * M);
            vx[r] = vext<UH>(a.ph, ix, iy, (unsigned)w, base + iy * (unsigned)a.ph.nz + (unsigned)w, it);
        }
    }
}

// One plane pass (MODE: PlaneMode) over every plane of the batch: workgroup
// blockIdx.x takes plane (b, ix) = (blockIdx.x / nx, blockIdx.x % nx).  The
// physics of the middle runs on the registers of the inverse y transform's
// last stage, and the kinetic factor and the store on those of the forward y
// transform's last stage.  UH: the user hook file (per-element ops).
template <typename T, int L, int MODE, bool UH>
__global__ __launch_bounds__(plane_threads<L>()) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && L >= 64 ? 4 : 2)))
void plane_kernel(const PassArgs<T> a) {
    using PL = PlaneLast<T, L>;
    constexpr int kPlaneThreads = plane_threads<L>();
    constexpr int P = L + 1;
    constexpr int NPT = L * L;
    constexpr int PE = NPT / kPlaneThreads;  // elements per thread in the row-major loops
    static_assert(NPT % kPlaneThreads == 0, "plane size");
    constexpr int R = PL::R, M = PL::M, PER = PL::PER;
    constexpr int NS = plan_of(L, false, sizeof(T) == 4).ns;
    __shared__ PlaneLds<T, L> S;
    for (int i = threadIdx.x; i < L; i += kPlaneThreads) {
        if constexpr (sizeof(T) == 4) {
            const cplx<float> h = a.tw[i], lo = a.twlo[i];
            S.tw[i].t = make_float4(h.x, h.y, lo.x, lo.y);
            // hi + lo float pairs of the FP64 kinetic factors
            const cplx<double> ky = a.Ky[i], kz = a.Kz[i];
            const float yh = (float)ky.x, yi = (float)ky.y, zh = (float)kz.x, zi = (float)kz.y;
            S.ky[i].t = make_float4(yh, yi, (float)(ky.x - (double)yh), (float)(ky.y - (double)yi));
            S.kz[i].t = make_float4(zh, zi, (float)(kz.x - (double)zh), (float)(kz.y - (double)zi));
        } else {
            S.tw[i].t = a.tw[i];
            S.ky[i].t = a.Ky[i];
            S.kz[i].t = a.Kz[i];
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
    const T cx = a.ph.c.x, cy = a.ph.c.y;
    double acc[3] = {0.0, 0.0, 0.0};
    cplx<T> v[PER][R];

    if constexpr (MODE == PL_OPEN) {
        // psi0 with the step's Vstep1 (orginal/gpe_engine.cu:723-772), row-major
        cplx<T> u[PE];
#pragma unroll
        for (int k = 0; k < PE; k++) u[k] = a.psiA[base + (unsigned)(threadIdx.x + k * kPlaneThreads)];
        if constexpr (!UH) {
            T vx[PE], gv[PE];
#pragma unroll
            for (int k = 0; k < PE; k++) {
                const int e = threadIdx.x + k * kPlaneThreads;
                const unsigned iy = (unsigned)(e / L), iz = (unsigned)(e % L);
                vx[k] = sep ? vscale * (a.ph.vtab[ix] + a.ph.vtab[nx + iy] + a.ph.vtab[nx + a.ph.ny + iz])
                            : vext(a.ph, ix, iy, iz, base + (unsigned)e, it);
            }
            dedfdn_chunk(a.ph, u, gv);
#pragma unroll
            for (int k = 0; k < PE; k++) vx[k] = vx[k] + gv[k];
            mul_exp_chunk(u, vx, cx, cy);
        } else {
#pragma unroll
            for (int k = 0; k < PE; k++) {
                const int e = threadIdx.x + k * kPlaneThreads;
                const unsigned iy = (unsigned)(e / L), iz = (unsigned)(e % L), off = base + (unsigned)e;
                const ElemIn<T> in = plane_elem<UH>(a, ix, iy, iz, off, it, sep, vscale, false, false);
                u[k] = load_op_m<LOAD_PSI0, UH>(a, in, u[k], acc);
            }
        }
#pragma unroll
        for (int k = 0; k < PE; k++) {
            const int e = threadIdx.x + k * kPlaneThreads;
            S.x[(e / L) * P + e % L] = u[k];
        }
        __syncthreads();
    } else {
        // side inputs of the physics at this thread's last-stage elements, loaded early
        cplx<T> side_c[MODE == PL_VSTEP2 ? PER : 1][MODE == PL_VSTEP2 ? R : 1];
        T side_r[MODE == PL_VSTEP2 ? 1 : PER][MODE == PL_VSTEP2 ? 1 : R];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int b = threadIdx.x + q * kPlaneThreads;
            if ((PL::NB % kPlaneThreads) == 0 || b < PL::NB) {
                const int w = b % L, j = b / L;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const unsigned off = base + (unsigned)((j + r * M) * L + w);
                    if constexpr (MODE == PL_VSTEP2) side_c[q][r] = ld_nt(a.psiA + off);
                    else side_r[q][r] = ld_nt(a.vbuf + off);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PE; k++) {
            const int e = threadIdx.x + k * kPlaneThreads;
            S.x[(e / L) * P + e % L] = a.src[base + (unsigned)e];
        }
        __syncthreads();
        plane_stages<T, L, P, 1, +1, 0, NS>(S.x, S.tw);      // inverse z
        plane_stages<T, L, 1, P, +1, 0, NS - 1>(S.x, S.tw);  // inverse y but its last stage
        plane_last_y<T, L, +1>(S.x, S.tw, v);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int b = threadIdx.x + q * kPlaneThreads;
            if (!((PL::NB % kPlaneThreads) == 0 || b < PL::NB)) continue;
            const int w = b % L, j = b / L;
            T vx[R], gv[R], lv[R];
            if constexpr (!UH) plane_vext<false, T, R, M>(a, ix, w, j, base, it, sep, vscale, vx);
            if constexpr (MODE == PL_VSTEP2) {
                // predictor end and corrector potential (orginal/gpe_engine.cu:774-801)
                if constexpr (!UH) {
                    cplx<T> p0[R];
                    T V0[R];
#pragma unroll
                    for (int r = 0; r < R; r++) p0[r] = side_c[q][r];
                    dedfdn_chunk(a.ph, p0, gv);
#pragma unroll
                    for (int r = 0; r < R; r++) V0[r] = vx[r] + gv[r];
                    mul_exp_chunk(v[q], V0, cx, cy);  // phi3
                    dedfdn_chunk(a.ph, v[q], gv);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        lv[r] = (T)0.5 * (V0[r] + vx[r] + gv[r]);
                        st_nt(a.vbuf + base + (unsigned)((j + r * M) * L + w), lv[r]);
                    }
                    mul_exp_chunk(p0, lv, cx, cy);
#pragma unroll
                    for (int r = 0; r < R; r++) v[q][r] = p0[r];
                } else {
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const unsigned iy = (unsigned)(j + r * M), off = base + iy * (unsigned)L + (unsigned)w;
                        ElemIn<T> in = plane_elem<UH>(a, ix, iy, (unsigned)w, off, it, sep, vscale, false, false);
                        in.aux = side_c[q][r];
                        v[q][r] = mid_op_m<MID_VSTEP2, UH>(a, in, v[q][r], acc);
                    }
                }
            } else {
                // corrector end (orginal/gpe_engine.cu:870-878), then the next step's Vstep1
                if constexpr (!UH) {
#pragma unroll
                    for (int r = 0; r < R; r++) lv[r] = side_r[q][r];
                    mul_exp_chunk(v[q], lv, cx, cy);  // psi1
#pragma unroll
                    for (int r = 0; r < R; r++) a.psiA[base + (unsigned)((j + r * M) * L + w)] = v[q][r];
                    if constexpr (MODE == PL_VSTEP3_NEXT) {
                        dedfdn_chunk(a.ph, v[q], gv);
#pragma unroll
                        for (int r = 0; r < R; r++) lv[r] = vx[r] + gv[r];
                        mul_exp_chunk(v[q], lv, cx, cy);
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const unsigned iy = (unsigned)(j + r * M), off = base + iy * (unsigned)L + (unsigned)w;
                        ElemIn<T> in = plane_elem<UH>(a, ix, iy, (unsigned)w, off, it, sep, vscale, false, false);
                        in.raux = side_r[q][r];
                        v[q][r] = mid_op_m<MID_VSTEP3, UH>(a, in, v[q][r], acc);
                    }
                }
            }
            if constexpr (MODE != PL_VSTEP3_LAST) {
#pragma unroll
                for (int r = 0; r < R; r++) S.x[w + (j + r * M) * P] = v[q][r];
            }
        }
        if constexpr (MODE == PL_VSTEP3_LAST) return;  // psi1 is in psi (A)
        __syncthreads();
    }
    plane_stages<T, L, P, 1, -1, 0, NS>(S.x, S.tw);      // forward z
    plane_stages<T, L, 1, P, -1, 0, NS - 1>(S.x, S.tw);  // forward y but its last stage
    plane_last_y<T, L, -1>(S.x, S.tw, v);
    // K_y K_z / N of the split kinetic factor from the LDS tables (per-point
    // loads of a K_y K_z table from global memory had cost a plane pass
    // 11-14%, measured), and the store
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = threadIdx.x + q * kPlaneThreads;
        if ((PL::NB % kPlaneThreads) == 0 || b < PL::NB) {
            const int w = b % L, j = b / L;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const unsigned iy = (unsigned)(j + r * M);
                a.dst[base + iy * (unsigned)L + (unsigned)w] = plane_kin<T, L>(S, iy, (unsigned)w, v[q][r]);
            }
        }
    }
}

}  // namespace gpe
