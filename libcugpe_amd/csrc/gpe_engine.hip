// Host side of the MI355X split-step GPE engine and its C ABI
// (include/gpe_engine.h).
//
// One split step of orginal/gpe_engine.cu:948-1013 is issued as a chain of
// fused FFT axis passes (gpe_fft.hpp) on the engine's own HIP stream:
//
//   open   z: Vstep1 (+imprint/rescale) then forward z          psi  -> B
//   step   y fwd | x fwd * K * x inv | y inv                       B
//          z: inv z, Vstep2 (predictor end, corrector potential
//             V1 stored), fwd z of modify(psi) W                B,psi -> psi
//          y fwd | x fwd * K * x inv | y inv                      psi
//          z: inv z, Vstep3 (x W), and the next step's Vstep1 +
//             fwd z when another step follows                  psi  -> psi, B
//
// so a real-time step reads and writes the state about 8 times instead of
// the reference's 4 cuFFT transforms plus 5 pointwise kernels.  Imaginary
// time splits the Vstep2 pass around the FP64 norm (finalised on the device,
// no host round trip), and quantum friction adds one fused FFT round trip of
// psi per step (the reference evaluates the identical potential twice,
// orginal/gpe_engine.cu:953 and 987).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#define GPE_NO_CPLX  // the engine has its own complex template gpe::cplx
#include "gpe_engine.h"
#include "gpe_kernels.hpp"

using namespace gpe;

namespace {

struct EventPair {
    hipEvent_t a, b;
};

}  // namespace

struct gpe_engine_s {
    int nx = 0, ny = 0, nz = 0;
    size_t n = 0;
    int precision = GPE_PRECISION_C128;
    int interactions = 0, vext = 0, gpe_for = 0;
    double gamma = 2.0;
    double alpha = 0, beta = 0, dt = 0, npart = 0, t0 = 0;
    unsigned it = 0;
    double qf = 0.0;  // stored divided by gamma N (orginal/gpe_engine.cu:642)
    double up[kMaxUserParams] = {0};
    double vx0 = 0, vy0 = 0;
    int vQ = 1;
    bool vortex_set = false, phase_set = false, varr_set = false;

    std::vector<double> kx, ky, kz;
    std::vector<std::complex<double>> Kx, Ky, Kz;
    std::complex<double> c;

    hipStream_t stream = nullptr;
    void *A = nullptr, *B = nullptr, *C = nullptr;  // complex state arrays
    void *Vb = nullptr, *Vq = nullptr, *varr = nullptr;  // real arrays
    void *tw[3] = {nullptr, nullptr, nullptr};        // per-axis twiddles
    void *twlo[3] = {nullptr, nullptr, nullptr};      // complex64: low parts of the twiddles
    void *Ktab = nullptr, *ktab = nullptr;           // kinetic factors / wave numbers
    void *Khl = nullptr;                             // complex64: K_x and K_y K_z as hi + lo float pairs
    void *vtab = nullptr;                            // per-axis terms of separable potentials
    double *k64 = nullptr;
    double *partials = nullptr;
    size_t partials_len = 0;
    double *scal = nullptr;  // device scalars
    double *rout = nullptr;  // FP64 real output (density / currents / phase)
    double *phase = nullptr; // enforced phase
    int8_t *wind = nullptr;

    // profiling
    int prof_class = -1;
    std::vector<EventPair> prof_ev;
    size_t prof_used = 0;
};

namespace {

constexpr int kSlotPot = 3;   // scal[3] eext, scal[4] eint
constexpr int kSlotKin = 5;   // scal[5] ekin
constexpr int kSlotNorm = 6;  // scal[6] norm sum (query)
constexpr int kSlotCount = 7; // scal[7] vortex count

#define HIP_TRY(cmd)                                   \
    do {                                               \
        hipError_t err_ = (cmd);                       \
        if (err_ != hipSuccess) return (int)err_;      \
    } while (0)

inline size_t esize(const gpe_engine_s* e) { return e->precision == GPE_PRECISION_C64 ? 4 : 8; }

// exp(-2 pi i m / n) with the angle reduced to the first octant in integer arithmetic
void unit_root(long m, long n, double* re, double* im) {
    m %= n;
    if (m < 0) m += n;
    const long o = (8 * m) / n;
    const long num = 8 * m - o * n;
    const double phi = (M_PI / 4.0) * ((double)num / (double)n);
    const double phc = (M_PI / 4.0) * ((double)(n - num) / (double)n);
    double c, s;
    switch (o) {
    case 0: c = std::cos(phi); s = std::sin(phi); break;
    case 1: c = std::sin(phc); s = std::cos(phc); break;
    case 2: c = -std::sin(phi); s = std::cos(phi); break;
    case 3: c = -std::cos(phc); s = std::sin(phc); break;
    case 4: c = -std::cos(phi); s = -std::sin(phi); break;
    case 5: c = -std::sin(phc); s = -std::cos(phc); break;
    case 6: c = std::sin(phi); s = -std::cos(phi); break;
    default: c = std::cos(phc); s = -std::sin(phc); break;
    }
    *re = c;
    *im = -s;
}

// k_a(i) = 2 pi i/n for i < n/2, 2 pi (i - n)/n otherwise (orginal/gpe_engine.cu:446-471)
void fill_k(std::vector<double>& k, int n) {
    k.resize(n);
    int i, j;
    for (i = 0; i <= n / 2 - 1; i++) k[i] = 2. * M_PI / n * (double)i;
    j = -i;
    for (i = n / 2; i < n; i++) { k[i] = 2. * M_PI / n * (double)j; j++; }
}

// step coefficient and kinetic factors (orginal/gpe_engine.cu:476-517, 572-610)
void build_kinetic(gpe_engine_s* e) {
    const std::complex<double> c1(e->gamma * 0.5 * e->dt, 0.0), c2(-1.0 * e->beta, e->alpha);
    e->c = c1 / c2;
    const double g2 = e->gamma * e->gamma;
    auto fill = [&](std::vector<std::complex<double>>& K, const std::vector<double>& k, double div) {
        K.resize(k.size());
        for (size_t i = 0; i < k.size(); i++) K[i] = std::exp(e->c * k[i] * k[i] / g2) / div;
    };
    fill(e->Kx, e->kx, 1.0);
    fill(e->Ky, e->ky, 1.0);
    fill(e->Kz, e->kz, (double)e->n);
}

// Kinetic factors in FP64; for complex64 also K_x and the products K_y K_z
// per (iy, iz), each as a float pair hi + lo with hi + lo = K to ~1e-15.
int upload_kinetic(gpe_engine_s* e) {
    std::vector<cplx<double>> K;
    for (auto& v : e->Kx) K.push_back(mk(v.real(), v.imag()));
    for (auto& v : e->Ky) K.push_back(mk(v.real(), v.imag()));
    for (auto& v : e->Kz) K.push_back(mk(v.real(), v.imag()));
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(e->Ktab, K.data(), K.size() * sizeof(cplx<double>), hipMemcpyHostToDevice));
    if (e->Khl) {
        std::vector<cplx<float>> h;
        h.reserve(2 * ((size_t)e->nx + (size_t)e->ny * e->nz));
        auto split = [&](std::complex<double> v) {
            const float hr = (float)v.real(), hi = (float)v.imag();
            h.push_back(mk(hr, hi));
            h.push_back(mk((float)(v.real() - (double)hr), (float)(v.imag() - (double)hi)));
        };
        for (auto& v : e->Kx) split(v);
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) split(e->Ky[iy] * e->Kz[iz]);
        HIP_TRY(hipMemcpy(e->Khl, h.data(), h.size() * sizeof(cplx<float>), hipMemcpyHostToDevice));
    }
    return 0;
}

template <typename T>
int upload_static_tables(gpe_engine_s* e) {
    const int dims[3] = {e->nx, e->ny, e->nz};
    for (int a = 0; a < 3; a++) {
        std::vector<cplx<T>> tw(dims[a]), lo(dims[a]);
        for (int m = 0; m < dims[a]; m++) {
            double re, im;
            unit_root(m, dims[a], &re, &im);
            tw[m] = mk((T)re, (T)im);
            lo[m] = mk((T)(re - (double)tw[m].x), (T)(im - (double)tw[m].y));
        }
        HIP_TRY(hipMemcpy(e->tw[a], tw.data(), tw.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
        if (e->twlo[a])
            HIP_TRY(hipMemcpy(e->twlo[a], lo.data(), lo.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
    }
    std::vector<T> k;
    std::vector<double> k64;
    for (const auto* v : {&e->kx, &e->ky, &e->kz})
        for (double x : *v) {
            k.push_back((T)x);
            k64.push_back(x);
        }
    HIP_TRY(hipMemcpy(e->ktab, k.data(), k.size() * sizeof(T), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->k64, k64.data(), k64.size() * sizeof(double), hipMemcpyHostToDevice));
    return upload_kinetic(e);
}

int upload_kinetic_any(gpe_engine_s* e) { return upload_kinetic(e); }

// Per-axis terms of the separable external potentials, so that kernels add
// three table entries instead of evaluating the hook:
//   harmonic  0.5 x^2 w_a^2, x = i - n/2        (orginal/gpe_user_defined.h:79-101)
//   cosine    cos(2 pi P i / n)                  (multiplied by A in the kernel)
template <typename T>
int upload_potential_tables(gpe_engine_s* e) {
    std::vector<T> v;
    const int dims[3] = {e->nx, e->ny, e->nz};
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < dims[a]; i++) {
            double t = 0.0;
            if (e->vext == VEXT_HARMONIC) {
                const double x = (double)i - 1.0 * (dims[a] / 2);
                t = 0.5 * x * x * e->up[a] * e->up[a];
            } else if (e->vext == VEXT_COSINE) {
                const double tp = 2.0 * M_PI * e->up[1];
                t = std::cos(tp * (double)i / (double)dims[a]);
            }
            v.push_back((T)t);
        }
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(e->vtab, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

int upload_potential_any(gpe_engine_s* e) {
    return e->precision == GPE_PRECISION_C64 ? upload_potential_tables<float>(e) : upload_potential_tables<double>(e);
}

int axis_len(const gpe_engine_s* e, int axis) { return axis == 0 ? e->nx : axis == 1 ? e->ny : e->nz; }

template <typename T>
PassArgs<T> base_args(gpe_engine_s* e) {
    PassArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.psiA = (cplx<T>*)e->A;
    a.vbuf = (T*)e->Vb;
    a.vq = (T*)e->Vq;
    a.rout = e->rout;
    const cplx<double>* K = (const cplx<double>*)e->Ktab;
    a.Kx = K;
    a.Ky = K + e->nx;
    a.Kz = K + e->nx + e->ny;
    const T* k = (const T*)e->ktab;
    a.kx = k;
    a.ky = k + e->nx;
    a.kz = k + e->nx + e->ny;
    if constexpr (sizeof(T) == 4) {
        a.kx_hl = (const cplx<float>*)e->Khl;
        a.kyz_hl = e->Khl ? (const cplx<float>*)e->Khl + 2 * e->nx : nullptr;
    }
    a.kx64 = e->k64;
    a.ky64 = e->k64 + e->nx;
    a.kz64 = e->k64 + e->nx + e->ny;
    a.scal = e->scal;
    a.partials = e->partials;
    for (int i = 0; i < kMaxUserParams; i++) a.up64[i] = e->up[i];
    Phys<T>& p = a.ph;
    p.nx = e->nx;
    p.ny = e->ny;
    p.nz = e->nz;
    p.interactions = e->interactions;
    p.vext = e->vext;
    p.it = e->it;
    p.gamma = (T)e->gamma;
    p.c = mk((T)e->c.real(), (T)e->c.imag());
    for (int i = 0; i < kMaxUserParams; i++) p.up[i] = (T)e->up[i];
    p.varr = (const T*)e->varr;
    p.vtab = (const T*)e->vtab;
    p.vx0 = (T)e->vx0;
    p.vy0 = (T)e->vy0;
    p.vQ = (T)e->vQ;
    return a;
}

struct PassSpec {
    int axis;
    const void* src;
    void* dst;
    int dir1, dir2;
    int load, mid, store, flags, nred;
};

// launches one pass; returns the workgroup count (for the reduction finalize)
template <typename T>
int run_pass(gpe_engine_s* e, PassArgs<T> a, const PassSpec& s) {
    a.axis = s.axis;
    const int N = axis_len(e, s.axis);
    a.M = (unsigned)(e->n / (size_t)N);
    a.n1 = (unsigned)e->n;
    a.src = (const cplx<T>*)s.src;
    a.dst = (cplx<T>*)s.dst;
    a.tw = (const cplx<T>*)e->tw[s.axis];
    a.twlo = (const cplx<T>*)e->twlo[s.axis];
    a.dir1 = s.dir1;
    a.dir2 = s.dir2;
    a.load_mode = s.load;
    a.mid_mode = s.mid;
    a.store_mode = s.store;
    a.flags = s.flags;
    a.nred = s.nred;
    // profiling classes: 0 strided (x and y) passes, 1 z passes, 2 y passes, 3 x passes
    const bool match = (e->prof_class == 0 && s.axis != 2) || (e->prof_class == 1 && s.axis == 2) ||
                       (e->prof_class == 2 && s.axis == 1) || (e->prof_class == 3 && s.axis == 0);
    const bool prof = match && (e->prof_used < e->prof_ev.size());
    if (prof) (void)hipEventRecord(e->prof_ev[e->prof_used].a, e->stream);
    const unsigned blocks = launch_fft_pass<T>(N, a, e->stream);
    if (prof) {
        (void)hipEventRecord(e->prof_ev[e->prof_used].b, e->stream);
        e->prof_used++;
    }
    if (blocks == 0) return -1;
    return (int)blocks;
}

void finalize(gpe_engine_s* e, int nb, int nq, int mode, int slot, double div) {
    launch_finalize(e->partials, nb, nq, e->scal, mode, slot, e->npart, e->dt, div, e->stream);
}

template <typename T>
int normalize_t(gpe_engine_s* e) {
    const int nb = launch_norm_partials<T>((const cplx<T>*)e->A, (unsigned)e->n, e->gamma, e->partials, e->stream);
    finalize(e, nb, 1, FIN_NORM, 0, 1.0);
    launch_scale<T>((cplx<T>*)e->A, (unsigned)e->n, e->scal, e->stream);
    HIP_TRY(hipGetLastError());
    return 0;
}

// nt split steps (orginal/gpe_engine.cu:942-1016, dev/gpe_engine.cu:1030-1202)
// pin: 0 none, 1 vortex imprint, 2 imprint where |psi| > 1e-15, 3 stored phase;
// mid_pin: re-applied after the predictor normalisation in imaginary time
template <typename T>
int evolve_t(gpe_engine_s* e, int nt, double* mu, int pin, bool mid_pin) {
    if (nt <= 0) return 0;
    const bool ite = (e->beta != 0.0);
    const bool qf = (e->qf != 0.0);
    const bool vortex = (pin != 0);
    const int imp = vortex ? F_IMPRINT : 0;
    const int fqf = qf ? F_QF : 0;
    void *A = e->A, *B = e->B, *C = e->C;
    const int F = -1, I = +1;

    PassArgs<T> a = base_args<T>(e);
    a.kfac = (T)e->qf;
    a.pin = pin;
    a.phase = e->phase;
    if (!qf)
        run_pass<T>(e, a, {2, nullptr, B, F, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN, F_VSTEP1 | imp, 0});
    else
        run_pass<T>(e, a, {2, nullptr, C, F, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN, imp, 0});

    for (int s = 0; s < nt; s++) {
        const bool last = (s == nt - 1);
        a.ph.it = e->it;
        if (qf) {
            run_pass<T>(e, a, {1, C, C, F, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
            run_pass<T>(e, a, {0, C, C, F, I, LOAD_PLAIN, MID_K2, STORE_PLAIN, 0, 0});
            run_pass<T>(e, a, {1, C, C, I, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
            run_pass<T>(e, a, {2, C, B, I, F, LOAD_PLAIN, MID_QF_VSTEP1, STORE_PLAIN, F_QF, 0});
        }
        // predictor kinetic step of phi1
        run_pass<T>(e, a, {1, B, B, F, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {0, B, B, F, I, LOAD_PLAIN, MID_EXPT, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {1, B, B, I, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        if (!ite) {
            run_pass<T>(e, a, {2, B, A, I, F, LOAD_PLAIN, MID_VSTEP2, STORE_PLAIN, fqf, 0});
        } else {
            const int nb = run_pass<T>(e, a, {2, B, B, I, 0, LOAD_PLAIN, MID_PART1, STORE_PLAIN, fqf, 1});
            finalize(e, nb, 1, FIN_NORM, 0, 1.0);
            run_pass<T>(e, a, {2, B, A, F, 0, LOAD_PART2, MID_NONE, STORE_PLAIN,
                               F_SCALE | fqf | (mid_pin ? F_IMPRINT_MID : 0), 0});
        }
        // corrector kinetic step of modify(psi0) W
        run_pass<T>(e, a, {1, A, A, F, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {0, A, A, F, I, LOAD_PLAIN, MID_EXPT, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {1, A, A, I, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        if (!ite) {
            if (!last) {
                const int nxt = (qf ? F_NEXT_PSI : F_NEXT_VSTEP1) | (vortex ? F_IMPRINT_NEXT : 0);
                run_pass<T>(e, a, {2, A, qf ? C : B, I, F, LOAD_PLAIN, MID_VSTEP3, STORE_PLAIN, nxt, 0});
            } else {
                run_pass<T>(e, a, {2, A, nullptr, I, 0, LOAD_PLAIN, MID_VSTEP3, STORE_NONE, 0, 0});
            }
        } else {
            const int nb = run_pass<T>(e, a, {2, A, nullptr, I, 0, LOAD_PLAIN, MID_VSTEP3, STORE_NONE, F_RED_NORM, 1});
            finalize(e, nb, 1, FIN_NORM, 0, 1.0);
            if (!last) {
                if (!qf)
                    run_pass<T>(e, a, {2, nullptr, B, F, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN,
                                       F_SCALE | F_VSTEP1 | imp, 0});
                else
                    run_pass<T>(e, a, {2, nullptr, C, F, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN, F_SCALE | imp, 0});
            } else {
                launch_scale<T>((cplx<T>*)A, (unsigned)e->n, e->scal, e->stream);
            }
        }
        e->it = e->it + 1;
    }
    HIP_TRY(hipGetLastError());
    if (ite && mu) {
        HIP_TRY(hipMemcpyAsync(mu, e->scal + 2, sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return 0;
}

// gpe_energy: orginal/gpe_engine.cu:1076-1113; the kinetic term by Parseval,
// sum_k |k|^2 |psi_k|^2 / (2 gamma N), after one forward transform
template <typename T>
int energy_t(gpe_engine_s* e, double* t, double* ekin, double* eint, double* eext) {
    *t = e->t0 + e->dt * e->it;
    if (e->beta == 0.0) {
        int r = normalize_t<T>(e);
        if (r) return r;
    }
    PassArgs<T> a = base_args<T>(e);
    int nb = run_pass<T>(e, a, {2, e->A, e->B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, F_RED_POT, 2});
    finalize(e, nb, 2, FIN_STORE, kSlotPot, 1.0);
    run_pass<T>(e, a, {1, e->B, e->B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    nb = run_pass<T>(e, a, {0, e->B, nullptr, -1, 0, LOAD_PLAIN, MID_KIN_SUM, STORE_NONE, 0, 1});
    finalize(e, nb, 1, FIN_STORE, kSlotKin, 2.0 * e->gamma * (double)e->n);
    HIP_TRY(hipGetLastError());
    double h[3];
    HIP_TRY(hipMemcpyAsync(h, e->scal + kSlotPot, 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *eext = h[0];
    *eint = h[1];
    *ekin = h[2];
    return 0;
}

int ensure_rout(gpe_engine_s* e) {
    if (!e->rout) HIP_TRY(hipMalloc(&e->rout, e->n * sizeof(double)));
    return 0;
}

// gpe_get_currents: orginal/gpe_engine.cu:1164-1212
template <typename T>
int currents_t(gpe_engine_s* e, double* js[3]) {
    int r = ensure_rout(e);
    if (r) return r;
    PassArgs<T> a = base_args<T>(e);
    a.kfac = (T)(e->gamma * (double)e->n);
    for (int ax = 0; ax < 3; ax++) {
        a.ka_axis = ax;
        run_pass<T>(e, a, {2, e->A, e->B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {1, e->B, e->B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {0, e->B, e->B, -1, 1, LOAD_PLAIN, MID_KA, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {1, e->B, e->B, 1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {2, e->B, nullptr, 1, 0, LOAD_PLAIN, MID_OVERLAP_RE, STORE_NONE, 0, 0});
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(js[ax], e->rout, e->n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

template <typename T>
int set_psi_t(gpe_engine_s* e, const cuCplx* psi) {
    if (sizeof(T) == sizeof(double)) {
        HIP_TRY(hipMemcpyAsync(e->A, psi, e->n * sizeof(cuCplx), hipMemcpyHostToDevice, e->stream));
    } else {
        std::vector<cplx<T>> tmp(e->n);
        for (size_t i = 0; i < e->n; i++) tmp[i] = mk((T)psi[i].x, (T)psi[i].y);
        HIP_TRY(hipMemcpyAsync(e->A, tmp.data(), e->n * sizeof(cplx<T>), hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

template <typename T>
int get_psi_t(gpe_engine_s* e, cuCplx* psi) {
    if (sizeof(T) == sizeof(double)) {
        HIP_TRY(hipMemcpyAsync(psi, e->A, e->n * sizeof(cuCplx), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    } else {
        std::vector<cplx<T>> tmp(e->n);
        HIP_TRY(hipMemcpyAsync(tmp.data(), e->A, e->n * sizeof(cplx<T>), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        for (size_t i = 0; i < e->n; i++) {
            psi[i].x = (double)tmp[i].x;
            psi[i].y = (double)tmp[i].y;
        }
    }
    return 0;
}

void free_all(gpe_engine_s* e) {
    void* ptrs[] = {e->A, e->B, e->C, e->Vb, e->Vq, e->varr, e->tw[0], e->tw[1], e->tw[2], e->twlo[0], e->twlo[1],
                    e->twlo[2], e->Ktab, e->ktab, e->Khl, e->vtab,
                    e->k64, e->partials, e->scal, e->rout, e->phase, e->wind};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& ev : e->prof_ev) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
}

int create_impl(gpe_engine_s** out, const gpe_config_t* cfg, double alpha, double beta, double dt, double npart) {
    if (!out || !cfg) return GPE_ERR_ARG;
    *out = nullptr;
    if (cfg->nx < 2 || cfg->ny < 2 || cfg->nz < 2) return GPE_ERR_LATTICE;
    if (!fft_length_supported(cfg->nx) || !fft_length_supported(cfg->ny) || !fft_length_supported(cfg->nz))
        return GPE_ERR_LATTICE;
    const size_t n = (size_t)cfg->nx * cfg->ny * cfg->nz;
    if (n >= (size_t)1 << 31) return GPE_ERR_LATTICE;
    if (cfg->precision != GPE_PRECISION_C64 && cfg->precision != GPE_PRECISION_C128) return GPE_ERR_ARG;
    if (cfg->gpe_for != 0 && cfg->gpe_for != PARTICLES && cfg->gpe_for != DIMERS) return GPE_ERR_GAMMA;
    if (cfg->vext < 0 || cfg->vext > 3) return GPE_ERR_ARG;
    if (cfg->interactions < -2 || cfg->interactions > 1) return GPE_ERR_ARG;

    gpe_engine_s* e = new gpe_engine_s();
    e->nx = cfg->nx;
    e->ny = cfg->ny;
    e->nz = cfg->nz;
    e->n = n;
    e->precision = cfg->precision;
    e->interactions = cfg->interactions;
    e->vext = cfg->vext;
    e->gpe_for = cfg->gpe_for;
    if (cfg->gpe_for == PARTICLES) e->gamma = 1.0;
    else if (cfg->gpe_for == DIMERS) e->gamma = 2.0;
    else e->gamma = (cfg->interactions == 0 || cfg->interactions == 1) ? 2.0 : 1.0;
    e->alpha = alpha;
    e->beta = beta;
    e->dt = dt;
    e->npart = npart;
    fill_k(e->kx, e->nx);
    fill_k(e->ky, e->ny);
    fill_k(e->kz, e->nz);
    build_kinetic(e);

    auto fail = [&](int code) {
        free_all(e);
        delete e;
        return code;
    };
    hipError_t err;
#define ALLOC(ptr, bytes)                               \
    err = hipMalloc((void**)&(ptr), (bytes));           \
    if (err != hipSuccess) return fail((int)err);
    err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (err != hipSuccess) return fail((int)err);
    const size_t es = esize(e);
    ALLOC(e->A, n * 2 * es);
    ALLOC(e->B, n * 2 * es);
    ALLOC(e->Vb, n * es);
    for (int a = 0; a < 3; a++) {
        ALLOC(e->tw[a], (size_t)axis_len(e, a) * 2 * es);
        if (e->precision == GPE_PRECISION_C64) ALLOC(e->twlo[a], (size_t)axis_len(e, a) * 2 * es);
    }
    ALLOC(e->Ktab, (size_t)(e->nx + e->ny + e->nz) * sizeof(cplx<double>));
    ALLOC(e->ktab, (size_t)(e->nx + e->ny + e->nz) * es);
    if (e->precision == GPE_PRECISION_C64)
        ALLOC(e->Khl, 2 * ((size_t)e->nx + (size_t)e->ny * e->nz) * sizeof(cplx<float>));
    ALLOC(e->vtab, (size_t)(e->nx + e->ny + e->nz) * es);
    ALLOC(e->k64, (size_t)(e->nx + e->ny + e->nz) * sizeof(double));
    // partial sums: the largest grid of any pass times three sums
    size_t maxb = kMaxReduceBlocks;
    for (int a = 0; a < 3; a++) {
        const int N = axis_len(e, a);
        const int W = (e->precision == GPE_PRECISION_C64) ? fft_tile_width<float>(N) : fft_tile_width<double>(N);
        maxb = std::max(maxb, (n / N + W - 1) / W);
    }
    e->partials_len = 3 * maxb;
    ALLOC(e->partials, e->partials_len * sizeof(double));
    ALLOC(e->scal, 16 * sizeof(double));
#undef ALLOC
    (void)hipMemset(e->A, 0, n * 2 * es);
    (void)hipMemset(e->scal, 0, 16 * sizeof(double));
    int r = (e->precision == GPE_PRECISION_C64) ? upload_static_tables<float>(e) : upload_static_tables<double>(e);
    if (!r) r = upload_potential_any(e);
    if (r) return fail(r);
    *out = e;
    return 0;
}

template <typename T>
int detect_t(gpe_engine_s* e, double rho_min, int8_t* w, long* count) {
    if (!e->wind) HIP_TRY(hipMalloc(&e->wind, 3 * e->n));
    const int nb = launch_windings<T>((const cplx<T>*)e->A, e->nx, e->ny, e->nz, rho_min, e->wind, e->partials,
                                      e->stream);
    finalize(e, nb, 1, FIN_STORE, kSlotCount, 1.0);
    HIP_TRY(hipGetLastError());
    double c = 0;
    HIP_TRY(hipMemcpyAsync(&c, e->scal + kSlotCount, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (w) HIP_TRY(hipMemcpyAsync(w, e->wind, 3 * e->n, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (count) *count = (long)c;
    return 0;
}

#define DISPATCH(e, fn, ...) ((e)->precision == GPE_PRECISION_C64 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

// ------------------------------------------------------ singleton state
struct LegacyConfig {
    bool configured = false;
    gpe_config_t cfg{0, 0, 0, 0, 0, 0, GPE_PRECISION_C128};
};
LegacyConfig g_cfg;
gpe_engine_s* g_eng = nullptr;
std::mutex g_mu;

}  // namespace

// =================================================================== C ABI
extern "C" {

int gpe_engine_create(gpe_engine_t* out, const gpe_config_t* cfg, double alpha, double beta, double dt,
                      double npart) {
    return create_impl(out, cfg, alpha, beta, dt, npart);
}

int gpe_engine_destroy(gpe_engine_t e) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    free_all(e);
    delete e;
    return 0;
}

int gpe_engine_change_alpha_beta(gpe_engine_t e, double alpha, double beta) {
    if (!e) return GPE_ERR_NO_ENGINE;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->alpha = alpha;
    e->beta = beta;
    build_kinetic(e);
    return upload_kinetic_any(e);
}

int gpe_engine_set_time(gpe_engine_t e, double t0) {
    if (!e) return GPE_ERR_NO_ENGINE;
    e->t0 = t0;
    e->it = 0;
    return 0;
}

int gpe_engine_set_user_params(gpe_engine_t e, int size, const double* params) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (size > MAX_USER_PARAMS || size > kMaxUserParams) return GPE_ERR_TOO_MANY_PARAMS;
    if (size < 0 || (size > 0 && !params)) return GPE_ERR_ARG;
    for (int i = 0; i < kMaxUserParams; i++) e->up[i] = (i < size) ? params[i] : 0.0;
    return upload_potential_any(e);
}

// orginal/gpe_engine.cu:637-662
int gpe_engine_set_quantum_friction_coeff(gpe_engine_t e, double qfcoeff) {
    if (!e) return GPE_ERR_NO_ENGINE;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (qfcoeff != 0.0) {
        e->qf = qfcoeff / (e->gamma * (double)e->n);
        const size_t es = esize(e);
        if (!e->C) HIP_TRY(hipMalloc(&e->C, e->n * 2 * es));
        if (!e->Vq) HIP_TRY(hipMalloc(&e->Vq, e->n * es));
    } else {
        e->qf = 0.0;
        if (e->C) (void)hipFree(e->C);
        if (e->Vq) (void)hipFree(e->Vq);
        e->C = nullptr;
        e->Vq = nullptr;
    }
    return 0;
}

// dev/gpe_engine.cu:465-476
int gpe_engine_set_vortex(gpe_engine_t e, double x0, double y0, int Q) {
    if (!e) return GPE_ERR_NO_ENGINE;
    e->vx0 = x0;
    e->vy0 = y0;
    e->vQ = Q;
    e->vortex_set = true;
    return 0;
}

int gpe_engine_set_external_potential(gpe_engine_t e, const double* v) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!v) return GPE_ERR_ARG;
    const size_t es = esize(e);
    if (!e->varr) HIP_TRY(hipMalloc(&e->varr, e->n * es));
    if (es == 8) {
        HIP_TRY(hipMemcpy(e->varr, v, e->n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(v, v + e->n);
        HIP_TRY(hipMemcpy(e->varr, tmp.data(), e->n * sizeof(float), hipMemcpyHostToDevice));
    }
    e->varr_set = true;
    return 0;
}

// orginal/gpe_engine.cu:699-717
int gpe_engine_set_psi(gpe_engine_t e, double t, const cuCplx* psi) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!psi) return GPE_ERR_ARG;
    e->t0 = t;
    e->it = 0;
    return DISPATCH(e, set_psi_t, e, psi);
}

int gpe_engine_get_psi(gpe_engine_t e, double* t, cuCplx* psi) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!psi || !t) return GPE_ERR_ARG;
    int r = DISPATCH(e, get_psi_t, e, psi);
    *t = e->t0 + e->dt * e->it;
    return r;
}

// orginal/gpe_engine.cu:684-697
int gpe_engine_normalize_psi(gpe_engine_t e) {
    if (!e) return GPE_ERR_NO_ENGINE;
    return DISPATCH(e, normalize_t, e);
}

// dev/gpe_engine.cu:679-689
int gpe_engine_get_norm(gpe_engine_t e, double* norm) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!norm) return GPE_ERR_ARG;
    const int nb = (e->precision == GPE_PRECISION_C64)
                       ? launch_norm_partials<float>((const cplx<float>*)e->A, (unsigned)e->n, e->gamma, e->partials, e->stream)
                       : launch_norm_partials<double>((const cplx<double>*)e->A, (unsigned)e->n, e->gamma, e->partials, e->stream);
    finalize(e, nb, 1, FIN_STORE, kSlotNorm, 1.0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(norm, e->scal + kSlotNorm, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

static int evolve_any(gpe_engine_t e, int nt, double* mu, int pin, bool mid_pin) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if ((pin == 1 || pin == 2) && !e->vortex_set) return GPE_ERR_VORTEX_NOT_SET;
    if (pin == 3 && !e->phase_set) return GPE_ERR_PHASE_NOT_SET;
    if (e->vext == VEXT_ARRAY && !e->varr_set) return GPE_ERR_ARG;
    return e->precision == GPE_PRECISION_C64 ? evolve_t<float>(e, nt, mu, pin, mid_pin)
                                             : evolve_t<double>(e, nt, mu, pin, mid_pin);
}

// orginal/gpe_engine.cu:942-1016, dev/gpe_engine.cu:1030-1109
int gpe_engine_evolve_qf(gpe_engine_t e, int nt, double* mu) { return evolve_any(e, nt, mu, 0, false); }
// dev/gpe_engine.cu:1115-1202: imprint at every step start and after the ITE predictor
int gpe_engine_evolve_vortex(gpe_engine_t e, int nt, double* mu) { return evolve_any(e, nt, mu, 1, true); }
// dev/gpe_engine.cu:1207-1294: the same with the imprint only where |psi| > 1e-15
int gpe_engine_evolve_vortex2(gpe_engine_t e, int nt, double* mu) { return evolve_any(e, nt, mu, 2, true); }
// dev/gpe_engine.cu:1299-1386: imprint at the step start only
int gpe_engine_evolve_vortex3(gpe_engine_t e, int nt, double* mu) { return evolve_any(e, nt, mu, 1, false); }
// dev/gpe_engine.cu:1390-1477: the stored phase enforced like the vortex imprint
int gpe_engine_evolve_enforced_phase(gpe_engine_t e, int nt, double* mu) { return evolve_any(e, nt, mu, 3, true); }

int gpe_engine_energy(gpe_engine_t e, double* t, double* ekin, double* eint, double* eext) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!t || !ekin || !eint || !eext) return GPE_ERR_ARG;
    if (e->vext == VEXT_ARRAY && !e->varr_set) return GPE_ERR_ARG;
    return DISPATCH(e, energy_t, e, t, ekin, eint, eext);
}

// orginal/gpe_engine.cu:1115-1123
int gpe_engine_get_density(gpe_engine_t e, double* t, double* density) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!t || !density) return GPE_ERR_ARG;
    int r = ensure_rout(e);
    if (r) return r;
    if (e->precision == GPE_PRECISION_C64)
        launch_density<float>((const cplx<float>*)e->A, (unsigned)e->n, e->gamma, e->rout, e->stream);
    else
        launch_density<double>((const cplx<double>*)e->A, (unsigned)e->n, e->gamma, e->rout, e->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(density, e->rout, e->n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *t = e->t0 + e->dt * e->it;
    return 0;
}

int gpe_engine_get_currents(gpe_engine_t e, double* t, double* jx, double* jy, double* jz) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!t || !jx || !jy || !jz) return GPE_ERR_ARG;
    *t = e->t0 + e->dt * e->it;
    double* js[3] = {jx, jy, jz};
    return DISPATCH(e, currents_t, e, js);
}

// dev/gpe_engine.cu:552-582
int gpe_engine_get_phase(gpe_engine_t e, double* h_phase) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!e->phase) HIP_TRY(hipMalloc(&e->phase, e->n * sizeof(double)));
    if (e->precision == GPE_PRECISION_C64)
        launch_phase<float>((const cplx<float>*)e->A, (unsigned)e->n, e->phase, e->stream);
    else
        launch_phase<double>((const cplx<double>*)e->A, (unsigned)e->n, e->phase, e->stream);
    HIP_TRY(hipGetLastError());
    if (h_phase) HIP_TRY(hipMemcpyAsync(h_phase, e->phase, e->n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->phase_set = true;
    return 0;
}

int gpe_engine_set_phase(gpe_engine_t e, const double* h_phase) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!h_phase) return GPE_ERR_ARG;
    if (!e->phase) HIP_TRY(hipMalloc(&e->phase, e->n * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(e->phase, h_phase, e->n * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->phase_set = true;
    return 0;
}

int gpe_engine_detect_vortices(gpe_engine_t e, double rho_min, int8_t* windings, long* count) {
    if (!e) return GPE_ERR_NO_ENGINE;
    return DISPATCH(e, detect_t, e, rho_min, windings, count);
}

int gpe_engine_fft(gpe_engine_t e, int sign) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (sign != -1 && sign != 1) return GPE_ERR_ARG;
    for (int ax = 2; ax >= 0; ax--) {
        if (e->precision == GPE_PRECISION_C64)
            run_pass<float>(e, base_args<float>(e), {ax, e->A, e->A, sign, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        else
            run_pass<double>(e, base_args<double>(e), {ax, e->A, e->A, sign, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int gpe_engine_get_time(gpe_engine_t e, double* t) {
    if (!e) return GPE_ERR_NO_ENGINE;
    *t = e->t0 + e->dt * e->it;
    return 0;
}

int gpe_engine_get_config(gpe_engine_t e, gpe_config_t* cfg) {
    if (!e) return GPE_ERR_NO_ENGINE;
    cfg->nx = e->nx;
    cfg->ny = e->ny;
    cfg->nz = e->nz;
    cfg->interactions = e->interactions;
    cfg->vext = e->vext;
    cfg->gpe_for = (e->gamma == 1.0) ? PARTICLES : DIMERS;
    cfg->precision = e->precision;
    return 0;
}

int gpe_engine_device_psi(gpe_engine_t e, void** ptr) {
    if (!e) return GPE_ERR_NO_ENGINE;
    *ptr = e->A;
    return 0;
}

int gpe_engine_stream(gpe_engine_t e, void** stream) {
    if (!e) return GPE_ERR_NO_ENGINE;
    *stream = (void*)e->stream;
    return 0;
}

int gpe_engine_synchronize(gpe_engine_t e) {
    if (!e) return GPE_ERR_NO_ENGINE;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int gpe_engine_profile(gpe_engine_t e, int axis_class, int enable) {
    if (!e) return GPE_ERR_NO_ENGINE;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (!enable) {
        e->prof_class = -1;
        return 0;
    }
    if (axis_class < 0 || axis_class > 3) return GPE_ERR_ARG;
    if (e->prof_ev.empty()) {
        e->prof_ev.resize(8192);
        for (auto& ev : e->prof_ev) {
            HIP_TRY(hipEventCreate(&ev.a));
            HIP_TRY(hipEventCreate(&ev.b));
        }
    }
    e->prof_used = 0;
    e->prof_class = axis_class;
    return 0;
}

int gpe_engine_profile_read(gpe_engine_t e, double* total_ms, long* launches) {
    if (!e) return GPE_ERR_NO_ENGINE;
    HIP_TRY(hipStreamSynchronize(e->stream));
    double tot = 0.0;
    for (size_t i = 0; i < e->prof_used; i++) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e->prof_ev[i].a, e->prof_ev[i].b));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (long)e->prof_used;
    return 0;
}

int gpe_engine_copy_benchmark(gpe_engine_t e, double* ms) {
    if (!e) return GPE_ERR_NO_ENGINE;
    const size_t bytes = e->n * 2 * esize(e);
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, e->stream));
    launch_copy(e->B, e->A, bytes, e->stream);
    HIP_TRY(hipEventRecord(b, e->stream));
    HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms) *ms = t;
    return 0;
}

const char* gpe_build_info(void) {
    return "libgpe_hip: gfx950, wave64, FFT lengths 8 12 16 24 32 48 64 96 128 192 256 512, complex64/complex128";
}

// dev/gpe_engine.h:158-195 (HIP-based code table)
const char* gpe_get_error_string(gpe_result_t res) {
    switch (res) {
    case GPE_SUCCESS: return "success";
    case GPE_ERR_TOO_MANY_PARAMS: return "too many user parameters";
    case GPE_ERR_NO_ENGINE: return "engine not created";
    case GPE_ERR_LATTICE: return "lattice not configured or length not supported";
    case GPE_ERR_ARG: return "invalid argument";
    case GPE_ERR_VORTEX_NOT_SET: return "vortex not set (gpe_set_vortex)";
    case GPE_ERR_PHASE_NOT_SET: return "phase not set (gpe_set_phase / gpe_get_phase)";
    case GPE_ERR_GAMMA: return "GPE_FOR must be PARTICLES or DIMERS";
    default: break;
    }
    if (res > 0) return hipGetErrorString((hipError_t)res);
    return "Unknown error!";
}

// --------------------------------------------------- singleton surface

int gpe_configure_lattice(int nx, int ny, int nz, int interactions, int vext, int gpe_for, int precision) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cfg.cfg.nx = nx;
    g_cfg.cfg.ny = ny;
    g_cfg.cfg.nz = nz;
    g_cfg.cfg.interactions = interactions;
    g_cfg.cfg.vext = vext;
    g_cfg.cfg.gpe_for = gpe_for;
    g_cfg.cfg.precision = precision ? precision : GPE_PRECISION_C128;
    g_cfg.configured = true;
    return 0;
}

void gpe_get_lattice(int* _nx, int* _ny, int* _nz) {
    *_nx = g_cfg.cfg.nx;
    *_ny = g_cfg.cfg.ny;
    *_nz = g_cfg.cfg.nz;
}

int gpe_create_engine(double alpha, double beta, double dt, double npart, int nthreads) {
    (void)nthreads;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_cfg.configured) return GPE_ERR_LATTICE;
    if (g_eng) {
        gpe_engine_destroy(g_eng);
        g_eng = nullptr;
    }
    return create_impl(&g_eng, &g_cfg.cfg, alpha, beta, dt, npart);
}

int gpe_destroy_engine(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_eng) return GPE_ERR_NO_ENGINE;
    int r = gpe_engine_destroy(g_eng);
    g_eng = nullptr;
    return r;
}

int gpe_change_alpha_beta(double alpha, double beta) { return gpe_engine_change_alpha_beta(g_eng, alpha, beta); }
int gpe_set_rte_evolution(void) { return gpe_engine_change_alpha_beta(g_eng, 1.0, 0.0); }
int gpe_set_ite_evolution(void) { return gpe_engine_change_alpha_beta(g_eng, 0.0, 1.0); }
int gpe_set_time(double t0) { return gpe_engine_set_time(g_eng, t0); }
int gpe_set_user_params(int size, double* params) { return gpe_engine_set_user_params(g_eng, size, params); }
int gpe_set_quantum_friction_coeff(double q) { return gpe_engine_set_quantum_friction_coeff(g_eng, q); }
int gpe_set_vortex(double x0, double y0, int8_t Q) { return gpe_engine_set_vortex(g_eng, x0, y0, (int)Q); }
int gpe_get_phase(double* h_phase) { return gpe_engine_get_phase(g_eng, h_phase); }
int gpe_set_phase(double* h_phase) { return gpe_engine_set_phase(g_eng, h_phase); }
int gpe_set_psi(double t, cuCplx* psi) { return gpe_engine_set_psi(g_eng, t, psi); }
int gpe_get_psi(double* t, cuCplx* psi) { return gpe_engine_get_psi(g_eng, t, psi); }
int gpe_get_density(double* t, double* d) { return gpe_engine_get_density(g_eng, t, d); }
int gpe_get_currents(double* t, double* jx, double* jy, double* jz) {
    return gpe_engine_get_currents(g_eng, t, jx, jy, jz);
}
int gpe_normalize_psi(void) { return gpe_engine_normalize_psi(g_eng); }
int gpe_get_norm(double* norm) { return gpe_engine_get_norm(g_eng, norm); }
int gpe_evolve(int nt) { return gpe_engine_evolve_qf(g_eng, nt, nullptr); }
int gpe_evolve_qf(int nt, double* mu) { return gpe_engine_evolve_qf(g_eng, nt, mu); }
int gpe_evolve_vortex(int nt, double* mu) { return gpe_engine_evolve_vortex(g_eng, nt, mu); }
int gpe_evolve_vortex2(int nt, double* mu) { return gpe_engine_evolve_vortex2(g_eng, nt, mu); }
int gpe_evolve_vortex3(int nt, double* mu) { return gpe_engine_evolve_vortex3(g_eng, nt, mu); }
int gpe_evolve_enforced_phase(int nt, double* mu) { return gpe_engine_evolve_enforced_phase(g_eng, nt, mu); }
int gpe_energy(double* t, double* ekin, double* eint, double* eext) {
    return gpe_engine_energy(g_eng, t, ekin, eint, eext);
}
int gpe_set_external_potential(const double* v) { return gpe_engine_set_external_potential(g_eng, v); }
int gpe_detect_vortices(double rho_min, int8_t* w, long* count) {
    return gpe_engine_detect_vortices(g_eng, rho_min, w, count);
}

int gpe_set_device(int device) {
    HIP_TRY(hipSetDevice(device));
    return 0;
}

int gpe_host_alloc(void** ptr, size_t bytes) {
    if (!ptr) return GPE_ERR_ARG;
    HIP_TRY(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return 0;
}

int gpe_host_free(void* ptr) {
    if (ptr) HIP_TRY(hipHostFree(ptr));
    return 0;
}

void gpe_print_potential_type(void) {
    static const char* names[] = {"HARMONIC POTENTIAL", "WOODS-SAXON TUBE", "COSINE LATTICE", "USER ARRAY"};
    int v = g_eng ? g_eng->vext : g_cfg.cfg.vext;
    printf("%s\n", (v >= 0 && v <= 3) ? names[v] : "UNKNOWN");
}

}  // extern "C"
