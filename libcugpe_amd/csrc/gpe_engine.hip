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
#include <array>
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
    size_t n = 0;        // points of one lattice
    size_t batch = 1;    // lattices of this shape stored one after another (ensembles)
    int variant = 0;     // kernel choice of the pass launcher (autotuned)
    int cus = 256;       // compute units of the device (persistent pass grids)
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
    std::vector<char> varr_slots;  // lattices of the batch whose VEXT 3 array has been set

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
    double *updev = nullptr;     // user parameters in FP64 per lattice [batch][kMaxUserParams] (user hooks)
    unsigned *itdev = nullptr;   // time index of the running step (user hooks)
    bool uses_hooks = false;     // INTERACTIONS or VEXT 100: the user hook file, time dependent
    double *partials = nullptr;
    size_t partials_len = 0;
    double *scal = nullptr;  // device scalars
    double *rout = nullptr;  // FP64 real output (density / currents / phase)
    double *phase = nullptr; // enforced phase
    int8_t *wind = nullptr;

    // adjoint of a real-time run (gpe_engine_adjoint_forward / _backward)
    void *adj_ck = nullptr, *adj_seg = nullptr;  // checkpoints, one segment of states
    size_t adj_ck_len = 0, adj_seg_len = 0;     // in states
    void *adj_lam = nullptr, *adj_phi = nullptr, *adj_lam0 = nullptr, *adj_tgt = nullptr, *adj_fin = nullptr;
    double *adj_vb = nullptr, *adj_grad = nullptr, *adj_vx = nullptr;
    int adj_nt = 0, adj_C = 0;
    unsigned adj_it0 = 0;
    bool adj_ready = false;
    unsigned adj_gen = 0;  // gen at the forward run: a parameter change since then voids the backward run
    double adj_fr = 0, adj_fi = 0;  // lambda(T) = (fr + i fi) psi_target

    // captured step graphs (open, middle step, last step) for one step plan;
    // gen counts parameter changes, a graph built at another gen is stale
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    int graph_pin = -1;
    bool graph_mid_pin = false;
    unsigned graph_gen = 0, gen = 1;
    bool use_graph = false;  // measured: direct launches are faster on one stream

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
    p.itp = nullptr;
    p.it_off = 0;
    p.modify = e->uses_hooks ? 1 : 0;
    p.uparam = e->updev;
    p.n1 = (unsigned)e->n;
    p.dt = e->dt;
    p.t0 = e->t0;
    p.npart = e->npart;
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
    a.M = (unsigned)(e->n * e->batch / (size_t)N);
    a.variant = e->variant;
    a.pipe_wgs = 2u * (unsigned)e->cus;
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

void drop_graphs(gpe_engine_s* e) {
    for (auto& g : e->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    e->graph_gen = 0;
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

// Step issue.  A run of nt split steps (orginal/gpe_engine.cu:942-1016,
// dev/gpe_engine.cu:1030-1202) is the open pass followed by nt steps; every
// step but the last ends in a pass that also opens the next one.
// pin: 0 none, 1 vortex imprint, 2 imprint where |psi| > 1e-15, 3 stored phase;
// mid_pin: re-applied after the predictor normalisation in imaginary time
struct StepPlan {
    int pin = 0;
    bool mid_pin = false;
};

template <typename T>
PassArgs<T> step_args(gpe_engine_s* e, const StepPlan& sp) {
    PassArgs<T> a = base_args<T>(e);
    a.kfac = (T)e->qf;
    a.pin = sp.pin;
    a.phase = e->phase;
    // user hooks read the step's time index from device memory, so that a
    // captured step graph stays valid for every step it is replayed for
    if (e->uses_hooks) a.ph.itp = e->itdev;
    return a;
}

// Launcher variant bit 5: the y-z plane plan (gpe_plane.hpp) where a plane
// of ny = nz points fits in one workgroup's LDS: four passes per real-time
// step (two plane passes, two x turn-arounds) instead of six.
constexpr int kVariantPlane = 32;
// Launcher variant bit 6: the alternative radix schedule (alt_plan) in the
// strided turn-arounds of the split kinetic factor.
constexpr int kVariantAltPlan = 64;

bool plane_eligible(const gpe_engine_s* e, int pin) {
    return e->ny == e->nz && plane_length_ok(e->ny) && e->beta == 0.0 && e->qf == 0.0 && pin == 0 &&
           e->interactions != INT_DIPOLAR;
}

bool plane_selected(const gpe_engine_s* e, int pin) { return (e->variant & kVariantPlane) && plane_eligible(e, pin); }

template <typename T>
void plane_run(gpe_engine_s* e, PassArgs<T> a, int mode, const void* src, void* dst, int flags) {
    a.src = (const cplx<T>*)src;
    a.dst = (cplx<T>*)dst;
    a.tw = (const cplx<T>*)e->tw[2];  // ny = nz: the y and z tables are the same
    a.twlo = (const cplx<T>*)e->twlo[2];
    a.n1 = (unsigned)e->n;
    a.flags = flags;
    (void)launch_plane<T>(a, e->ny, mode, (unsigned)(e->batch * (size_t)e->nx), e->uses_hooks, e->stream);
}

template <typename T>
void issue_open(gpe_engine_s* e, const StepPlan& sp) {
    if (plane_selected(e, sp.pin)) {
        plane_run<T>(e, step_args<T>(e, sp), PL_OPEN, nullptr, e->B, F_VSTEP1);
        return;
    }
    const bool qf = (e->qf != 0.0);
    const int imp = sp.pin ? F_IMPRINT : 0;
    PassArgs<T> a = step_args<T>(e, sp);
    if (!qf)
        run_pass<T>(e, a, {2, nullptr, e->B, -1, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN, F_VSTEP1 | imp, 0});
    else
        run_pass<T>(e, a, {2, nullptr, e->C, -1, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN, imp, 0});
}

// The kinetic factor of a step on a buffer whose z axis is already in k_z
// (orginal/gpe_engine.cu:963-966, FFT, exp(c k^2 / gamma^2) / N, inverse FFT).
// K = K_x K_y K_z factorises by axis and the 1-D transforms along x and y
// commute with each other and with the factors of the other axes, so the
// default plan is two turn-around passes: y (forward y, K_y K_z, inverse y)
// and x (forward x, K_x, inverse x).  Launcher variant bit 3 selects the
// unsplit plan: forward y, the x turn-around with K_x K_y K_z, inverse y.
template <typename T>
void kinetic_kz(gpe_engine_s* e, const PassArgs<T>& a, void* buf) {
    const int F = -1, I = +1;
    if (!(e->variant & 8)) {
        run_pass<T>(e, a, {1, buf, buf, F, I, LOAD_PLAIN, MID_EXPT_SEP, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {0, buf, buf, F, I, LOAD_PLAIN, MID_EXPT_SEP, STORE_PLAIN, 0, 0});
    } else {
        run_pass<T>(e, a, {1, buf, buf, F, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {0, buf, buf, F, I, LOAD_PLAIN, MID_EXPT, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {1, buf, buf, I, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    }
}

// one split step from time index e->it (the host counter is advanced by the caller)
template <typename T>
void issue_step(gpe_engine_s* e, const StepPlan& sp, bool last) {
    if (plane_selected(e, sp.pin)) {
        // x turn-around, plane pass, x turn-around, plane pass (gpe_plane.hpp)
        const PassArgs<T> a = step_args<T>(e, sp);
        void *A = e->A, *B = e->B;
        run_pass<T>(e, a, {0, B, B, -1, +1, LOAD_PLAIN, MID_EXPT_SEP, STORE_PLAIN, 0, 0});
        plane_run<T>(e, a, PL_VSTEP2, B, A, 0);
        run_pass<T>(e, a, {0, A, A, -1, +1, LOAD_PLAIN, MID_EXPT_SEP, STORE_PLAIN, 0, 0});
        if (!last) plane_run<T>(e, a, PL_VSTEP3_NEXT, A, B, F_NEXT_VSTEP1);
        else plane_run<T>(e, a, PL_VSTEP3_LAST, A, nullptr, 0);
        if (e->uses_hooks) launch_tick(e->itdev, e->stream);
        return;
    }
    const bool ite = (e->beta != 0.0);
    const bool qf = (e->qf != 0.0);
    const bool vortex = (sp.pin != 0);
    const int imp = vortex ? F_IMPRINT : 0;
    const int fqf = qf ? F_QF : 0;
    void *A = e->A, *B = e->B, *C = e->C;
    const int F = -1, I = +1;
    PassArgs<T> a = step_args<T>(e, sp);
    if (qf) {
        run_pass<T>(e, a, {1, C, C, F, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {0, C, C, F, I, LOAD_PLAIN, MID_K2, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {1, C, C, I, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {2, C, B, I, F, LOAD_PLAIN, MID_QF_VSTEP1, STORE_PLAIN, F_QF, 0});
    }
    // predictor kinetic step of phi1
    kinetic_kz<T>(e, a, B);
    if (!ite) {
        run_pass<T>(e, a, {2, B, A, I, F, LOAD_PLAIN, MID_VSTEP2, STORE_PLAIN, fqf, 0});
    } else {
        const int nb = run_pass<T>(e, a, {2, B, B, I, 0, LOAD_PLAIN, MID_PART1, STORE_PLAIN, fqf, 1});
        finalize(e, nb, 1, FIN_NORM, 0, 1.0);
        run_pass<T>(e, a, {2, B, A, F, 0, LOAD_PART2, MID_NONE, STORE_PLAIN,
                           F_SCALE | fqf | (sp.mid_pin ? F_IMPRINT_MID : 0), 0});
    }
    // corrector kinetic step of modify(psi0) W
    kinetic_kz<T>(e, a, A);
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
            PassArgs<T> an = a;
            an.ph.it_off = 1;  // the next step's Vstep1
            if (!qf)
                run_pass<T>(e, an, {2, nullptr, B, F, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN, F_SCALE | F_VSTEP1 | imp, 0});
            else
                run_pass<T>(e, an, {2, nullptr, C, F, 0, LOAD_PSI0, MID_NONE, STORE_PLAIN, F_SCALE | imp, 0});
        } else {
            launch_scale<T>((cplx<T>*)A, (unsigned)e->n, e->scal, e->stream);
        }
    }
    if (e->uses_hooks) launch_tick(e->itdev, e->stream);
}

// Launcher variant bit 4: whole real-time steps in one kernel per lattice
// (fused_steps_kernel), for cubes whose state fits in one workgroup's LDS.
// The bits below it still choose the passes of the runs the fused kernel does
// not cover (imaginary or complex time, quantum friction, pinning).
constexpr int kVariantFused = 16;

template <typename T>
bool fused_eligible(const gpe_engine_s* e, int pin) {
    return e->nx == e->ny && e->ny == e->nz && fused_length_ok<T>(e->nx) && e->beta == 0.0 && e->qf == 0.0 &&
           pin == 0 && e->interactions != INT_DIPOLAR;
}

template <typename T>
bool fused_selected(const gpe_engine_s* e, int pin) {
    return (e->variant & kVariantFused) && e->prof_class < 0 && fused_eligible<T>(e, pin);
}

// nt real-time steps of every lattice of the engine from time index e->it
// (the host counter is advanced by the caller), in launches of at most 256
// steps so that no single kernel runs for long
template <typename T>
int fused_run(gpe_engine_s* e, int nt, hipStream_t s) {
    const PassArgs<T> a = base_args<T>(e);
    FusedArgs<T> f;
    std::memset(&f, 0, sizeof(f));
    f.psi = (cplx<T>*)e->A;
    f.Kx = a.Kx;
    f.Ky = a.Ky;
    f.Kz = a.Kz;
    f.kx_hl = a.kx_hl;
    f.kyz_hl = a.kyz_hl;
    f.ph = a.ph;
    constexpr int kChunk = 256;
    for (int done = 0; done < nt; done += kChunk) {
        f.nt = std::min(kChunk, nt - done);
        f.it0 = e->it + (unsigned)done;
        if (launch_fused_steps<T>(f, e->nx, (unsigned)e->batch, e->uses_hooks, s)) return GPE_ERR_ARG;
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename T>
int evolve_t(gpe_engine_s* e, int nt, double* mu, int pin, bool mid_pin) {
    if (nt <= 0) return 0;
    if (fused_selected<T>(e, pin)) {
        if (int r = fused_run<T>(e, nt, e->stream)) return r;
        e->it = e->it + (unsigned)nt;
        return 0;
    }
    const bool ite = (e->beta != 0.0);
    // lattices of a batch share one stream of passes: imaginary time (one
    // norm per lattice) and a stored phase (one array) are single-lattice only
    if (e->batch > 1 && (ite || pin == 3)) return GPE_ERR_ARG;
    StepPlan sp;
    sp.pin = pin;
    sp.mid_pin = mid_pin;
    if (e->uses_hooks) launch_set_counter(e->itdev, e->it, e->stream);
    // With graphs on, runs of 3+ steps replay captured graphs: the passes'
    // arguments are the same every step (the user hooks' time index is read
    // from device memory, advanced by a kernel at the end of every step), so
    // one capture of the open pass, a middle step and the last step serves
    // every run until a parameter changes.  Not while profiling.
    if (e->use_graph && nt >= 3 && e->prof_class < 0 && e->batch == 1) {
        if (e->graph_gen != e->gen || e->graph_pin != pin || e->graph_mid_pin != mid_pin) {
            drop_graphs(e);
            for (int w = 0; w < 3; w++) {
                HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                if (w == 0) issue_open<T>(e, sp);
                else issue_step<T>(e, sp, w == 2);
                hipGraph_t g = nullptr;
                HIP_TRY(hipStreamEndCapture(e->stream, &g));
                const hipError_t err = hipGraphInstantiate(&e->graph[w], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIP_TRY(err);
            }
            e->graph_gen = e->gen;
            e->graph_pin = pin;
            e->graph_mid_pin = mid_pin;
        }
        HIP_TRY(hipGraphLaunch(e->graph[0], e->stream));
        for (int s = 0; s < nt - 1; s++) HIP_TRY(hipGraphLaunch(e->graph[1], e->stream));
        HIP_TRY(hipGraphLaunch(e->graph[2], e->stream));
        e->it = e->it + (unsigned)nt;
    } else {
        issue_open<T>(e, sp);
        for (int s = 0; s < nt; s++) {
            issue_step<T>(e, sp, s == nt - 1);
            e->it = e->it + 1;
        }
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
int set_psi_t(gpe_engine_s* e, const cuCplx* psi, size_t inst = 0) {
    cplx<T>* dst = (cplx<T>*)e->A + inst * e->n;
    if (sizeof(T) == sizeof(double)) {
        HIP_TRY(hipMemcpyAsync(dst, psi, e->n * sizeof(cuCplx), hipMemcpyHostToDevice, e->stream));
    } else {
        // complex128 values staged in the predictor array B (free between
        // calls), converted on the device: chunks of B's capacity
        const size_t cap = e->n * e->batch * sizeof(cplx<T>) / sizeof(cuCplx);
        for (size_t i0 = 0; i0 < e->n; i0 += cap) {
            const size_t m = std::min(cap, e->n - i0);
            HIP_TRY(hipMemcpyAsync(e->B, psi + i0, m * sizeof(cuCplx), hipMemcpyHostToDevice, e->stream));
            launch_c128_to_c64(e->B, dst + i0, m, e->stream);
            HIP_TRY(hipGetLastError());
        }
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

template <typename T>
int get_psi_t(gpe_engine_s* e, cuCplx* psi, size_t inst = 0) {
    const cplx<T>* src = (const cplx<T>*)e->A + inst * e->n;
    if (sizeof(T) == sizeof(double)) {
        HIP_TRY(hipMemcpyAsync(psi, src, e->n * sizeof(cuCplx), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    } else {
        // converted on the device into B, then copied out in chunks of B's capacity
        const size_t cap = e->n * e->batch * sizeof(cplx<T>) / sizeof(cuCplx);
        for (size_t i0 = 0; i0 < e->n; i0 += cap) {
            const size_t m = std::min(cap, e->n - i0);
            launch_c64_to_c128(src + i0, e->B, m, e->stream);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(psi + i0, e->B, m * sizeof(cuCplx), hipMemcpyDeviceToHost, e->stream));
        }
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return 0;
}

void free_all(gpe_engine_s* e) {
    drop_graphs(e);
    void* ptrs[] = {e->A, e->B, e->C, e->Vb, e->Vq, e->varr, e->tw[0], e->tw[1], e->tw[2], e->twlo[0], e->twlo[1],
                    e->twlo[2], e->Ktab, e->ktab, e->Khl, e->vtab,
                    e->k64, e->updev, e->itdev, e->partials, e->scal, e->rout, e->phase, e->wind, e->adj_ck, e->adj_seg,
                    e->adj_lam, e->adj_phi, e->adj_lam0, e->adj_tgt, e->adj_fin, e->adj_vb, e->adj_grad,
                    e->adj_vx};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& ev : e->prof_ev) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
}

// Launcher variant of a new engine: half-width tiles in the strided passes
// when the full ones would give the chip fewer than four workgroups per CU
// (measured in complex64: +21% steps/s at 64^3 and +12% at 128^3, while full
// tiles are 25% faster at 256^3).  The autotuner and set_variant override it.
static int default_variant(const gpe_engine_s* e) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    size_t most = 0;
    for (int a = 0; a < 2; a++) {  // x and y
        const int N = axis_len(e, a);
        const int W = (e->precision == GPE_PRECISION_C64) ? fft_lean_width<float>(N) : fft_lean_width<double>(N);
        if (W <= 0) return 0;
        most = std::max(most, (e->n * e->batch / (size_t)N + (size_t)W - 1) / (size_t)W);
    }
    int v = most < 4 * (size_t)cus ? 2 : 0;
    // the y-z plane plan for 32 x 32 and 48 x 48 planes (measured in complex64:
    // one 32^3 lattice 39k -> 51k steps/s, 48^3 28k -> 32k; 64 x 64 planes only
    // win in batches, where the autotuner finds them)
    if (e->ny == e->nz && (e->ny == 32 || e->ny == 48)) v |= kVariantPlane;
    return v;
}

int create_impl(gpe_engine_s** out, const gpe_config_t* cfg, double alpha, double beta, double dt, double npart,
                size_t batch = 1) {
    if (!out || !cfg) return GPE_ERR_ARG;
    *out = nullptr;
    if (cfg->nx < 2 || cfg->ny < 2 || cfg->nz < 2) return GPE_ERR_LATTICE;
    if (!fft_length_supported(cfg->nx) || !fft_length_supported(cfg->ny) || !fft_length_supported(cfg->nz))
        return GPE_ERR_LATTICE;
    const size_t n = (size_t)cfg->nx * cfg->ny * cfg->nz;
    if (batch < 1 || n * batch >= (size_t)1 << 31) return GPE_ERR_LATTICE;
    if (cfg->precision != GPE_PRECISION_C64 && cfg->precision != GPE_PRECISION_C128) return GPE_ERR_ARG;
    if (cfg->gpe_for != 0 && cfg->gpe_for != PARTICLES && cfg->gpe_for != DIMERS) return GPE_ERR_GAMMA;
    if ((cfg->vext < 0 || cfg->vext > 3) && cfg->vext != VEXT_USER) return GPE_ERR_ARG;
    if ((cfg->interactions < -2 || cfg->interactions > 1) && cfg->interactions != INT_USER) return GPE_ERR_ARG;

    gpe_engine_s* e = new gpe_engine_s();
    e->nx = cfg->nx;
    e->ny = cfg->ny;
    e->nz = cfg->nz;
    e->n = n;
    e->batch = batch;
    e->precision = cfg->precision;
    e->interactions = cfg->interactions;
    e->vext = cfg->vext;
    e->uses_hooks = (cfg->interactions == INT_USER || cfg->vext == VEXT_USER);
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
    const size_t nb = n * batch;
    ALLOC(e->A, nb * 2 * es);
    ALLOC(e->B, nb * 2 * es);
    ALLOC(e->Vb, nb * es);
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
        maxb = std::max(maxb, (nb / N + W - 1) / W);
    }
    e->partials_len = 3 * maxb;
    ALLOC(e->partials, e->partials_len * sizeof(double));
    ALLOC(e->scal, 16 * sizeof(double));
    ALLOC(e->updev, batch * kMaxUserParams * sizeof(double));
    ALLOC(e->itdev, 64);
#undef ALLOC
    (void)hipMemset(e->A, 0, nb * 2 * es);
    (void)hipMemset(e->scal, 0, 16 * sizeof(double));
    (void)hipMemset(e->updev, 0, batch * kMaxUserParams * sizeof(double));
    (void)hipMemset(e->itdev, 0, 64);
    int r = (e->precision == GPE_PRECISION_C64) ? upload_static_tables<float>(e) : upload_static_tables<double>(e);
    if (!r) r = upload_potential_any(e);
    if (r) return fail(r);
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            e->cus = cus;
    }
    e->variant = default_variant(e);
    *out = e;
    return 0;
}

// The FP64 parameter block the user hooks read as d_user_param, for lattice
// `slot` of the batch (every lattice when slot < 0).
int upload_user_params(gpe_engine_s* e, long slot, const double* up) {
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t b = 0; b < e->batch; b++)
        if (slot < 0 || b == (size_t)slot)
            HIP_TRY(hipMemcpy(e->updev + b * kMaxUserParams, up, kMaxUserParams * sizeof(double),
                              hipMemcpyHostToDevice));
    return 0;
}

// Every entry point that launches step or energy kernels calls this first:
// a VEXT 3 engine reads its potential array in every such kernel, so a
// missing array is an argument error raised on the host, before any launch.
int ready_to_launch(const gpe_engine_s* e) {
    if (e->vext == VEXT_ARRAY && !e->varr_set) return GPE_ERR_ARG;
    return 0;
}

// Cloud parameters of the state (dev/gpe_cloud_params.h:60-119) from one
// kernel over the device array instead of a host copy of psi:
//   out[0] the density at (nx/2, ny/2, nz/2)                        find_n0
//   out[1] the largest density                                      find_max_density
//   out[2..4] along the lines through the centre, the first distance
//   nx/2 - ix (ny/2 - iy, nz/2 - iz) with density below epsilon,
//   scanning ix = nx/2 down to 1; NaN when no point there is below  cloud_range_ix/iy/iz
template <typename T>
int cloud_t(gpe_engine_s* e, double epsilon, double* out) {
    if (int r = ensure_rout(e)) return r;
    const int nx = e->nx, ny = e->ny, nz = e->nz;
    const int nb = launch_cloud<T>((const cplx<T>*)e->A, nx, ny, nz, e->gamma, e->partials, e->rout, e->stream);
    HIP_TRY(hipGetLastError());
    std::vector<double> pm((size_t)nb), ln((size_t)(nx + ny + nz));
    HIP_TRY(hipMemcpyAsync(pm.data(), e->partials, pm.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(ln.data(), e->rout, ln.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    double m = 0.0;
    for (double v : pm) m = std::max(m, v);
    out[0] = ln[(size_t)(nx / 2)];
    out[1] = m;
    const int len[3] = {nx, ny, nz}, base[3] = {0, nx, nx + ny};
    for (int a = 0; a < 3; a++) {
        out[2 + a] = std::nan("");
        for (int i = len[a] / 2; i > 0; i--)
            if (ln[(size_t)(base[a] + i)] < epsilon) {
                out[2 + a] = (double)(len[a] / 2 - i);
                break;
            }
    }
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
    e->gen++;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->alpha = alpha;
    e->beta = beta;
    build_kinetic(e);
    return upload_kinetic_any(e);
}

int gpe_engine_set_time(gpe_engine_t e, double t0) {
    if (!e) return GPE_ERR_NO_ENGINE;
    e->gen++;
    e->t0 = t0;
    e->it = 0;
    return 0;
}

int gpe_engine_set_user_params(gpe_engine_t e, int size, const double* params) {
    if (!e) return GPE_ERR_NO_ENGINE;
    e->gen++;
    if (size > MAX_USER_PARAMS || size > kMaxUserParams) return GPE_ERR_TOO_MANY_PARAMS;
    if (size < 0 || (size > 0 && !params)) return GPE_ERR_ARG;
    for (int i = 0; i < kMaxUserParams; i++) e->up[i] = (i < size) ? params[i] : 0.0;
    int r = upload_user_params(e, -1, e->up);
    if (r) return r;
    return upload_potential_any(e);
}

// orginal/gpe_engine.cu:637-662
int gpe_engine_set_quantum_friction_coeff(gpe_engine_t e, double qfcoeff) {
    if (!e) return GPE_ERR_NO_ENGINE;
    e->gen++;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (qfcoeff != 0.0) {
        e->qf = qfcoeff / (e->gamma * (double)e->n);
        const size_t es = esize(e);
        if (!e->C) HIP_TRY(hipMalloc(&e->C, e->n * e->batch * 2 * es));
        if (!e->Vq) HIP_TRY(hipMalloc(&e->Vq, e->n * e->batch * es));
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
    e->gen++;
    e->vx0 = x0;
    e->vy0 = y0;
    e->vQ = Q;
    e->vortex_set = true;
    return 0;
}

// The VEXT 3 array of lattice `slot` of a batch (all lattices when slot < 0).
// The kernels index the array by the global offset, one copy per lattice.
static int set_varr_slots(gpe_engine_s* e, long slot, const double* v) {
    e->gen++;
    const size_t es = esize(e);
    if (!e->varr) HIP_TRY(hipMalloc(&e->varr, e->n * e->batch * es));
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<float> tmp;
    if (es == 4) tmp.assign(v, v + e->n);
    for (size_t b = 0; b < e->batch; b++) {
        if (slot >= 0 && b != (size_t)slot) continue;
        if (es == 8)
            HIP_TRY(hipMemcpy((double*)e->varr + b * e->n, v, e->n * sizeof(double), hipMemcpyHostToDevice));
        else
            HIP_TRY(hipMemcpy((float*)e->varr + b * e->n, tmp.data(), e->n * sizeof(float), hipMemcpyHostToDevice));
    }
    if (e->varr_slots.size() != e->batch) e->varr_slots.assign(e->batch, 0);
    for (size_t b = 0; b < e->batch; b++)
        if (slot < 0 || b == (size_t)slot) e->varr_slots[b] = 1;
    e->varr_set = std::all_of(e->varr_slots.begin(), e->varr_slots.end(), [](char c) { return c != 0; });
    return 0;
}

// The VEXT 3 array of every lattice of the engine from FP64 device memory,
// converted on the engine's stream (no host copy).
static int set_varr_device(gpe_engine_s* e, const double* d_v) {
    e->gen++;
    const size_t es = esize(e);
    if (!e->varr) HIP_TRY(hipMalloc(&e->varr, e->n * e->batch * es));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t b = 0; b < e->batch; b++) {
        if (es == 8)
            HIP_TRY(hipMemcpyAsync((double*)e->varr + b * e->n, d_v, e->n * sizeof(double), hipMemcpyDeviceToDevice,
                                   e->stream));
        else
            launch_f64_to_f32(d_v, (float*)e->varr + b * e->n, e->n, e->stream);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->varr_slots.assign(e->batch, 1);
    e->varr_set = true;
    return 0;
}

int gpe_engine_set_external_potential_device(gpe_engine_t e, const double* d_v) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!d_v || e->vext != VEXT_ARRAY) return GPE_ERR_ARG;
    return set_varr_device(e, d_v);
}

int gpe_engine_set_external_potential(gpe_engine_t e, const double* v) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!v) return GPE_ERR_ARG;
    return set_varr_slots(e, -1, v);
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
    if (int r = ready_to_launch(e)) return r;
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
    if (int r = ready_to_launch(e)) return r;
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
    e->gen++;
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
    e->gen++;
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

int gpe_engine_cloud_params(gpe_engine_t e, double epsilon, double* out) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!out) return GPE_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return DISPATCH(e, cloud_t, e, epsilon, out);
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
    case GPE_ERR_TOO_LARGE: return "lattice too large for gpe_detect_vortices_noalloc (nx*ny*nz < 2^31)";
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

int gpe_evolve_dipolar(int nt) { return gpe_engine_evolve_dipolar(g_eng, nt, nullptr); }
int gpe_evolve_dipolar_mu(int nt, double* chemical_potential) {
    return gpe_engine_evolve_dipolar(g_eng, nt, chemical_potential);
}
int gpe_energy_dipolar(double* t, double* edip) { return gpe_engine_energy_dipolar(g_eng, 1, t, edip); }
int gpe_energy_dipolar2(double* t, double* edip) { return gpe_engine_energy_dipolar(g_eng, 2, t, edip); }

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
    printf("%s\n", (v >= 0 && v <= 3) ? names[v] : (v == VEXT_USER ? "USER HOOK POTENTIAL" : "UNKNOWN"));
}

}  // extern "C"

// ====================================================================== dipolar
//
// INTERACTIONS -2 (dev/gpe_engine.cu:1490-1605): per step two potential half
// steps around the kinetic step, each with the convolution potential
// V_dip = Re F^-1[Vpp(k) F[rho]]/N of the current psi (the density goes to B
// as a complex array, one forward and one inverse FFT with the Vpp factor at
// the turn-around), then the ITE renormalisation.

namespace {

template <typename T>
void dipolar_potential_t(gpe_engine_s* e) {
    launch_density_c<T>((const cplx<T>*)e->A, (cplx<T>*)e->B, e->gamma, e->n, e->stream);
    PassArgs<T> a = base_args<T>(e);
    void* B = e->B;
    run_pass<T>(e, a, {2, B, B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {1, B, B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {0, B, B, -1, 1, LOAD_PLAIN, MID_VPP, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {1, B, B, 1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {2, B, B, 1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
}

template <typename T>
void kinetic_apply(gpe_engine_s* e, void* buf, bool adjoint);

template <typename T>
int evolve_dipolar_t(gpe_engine_s* e, int nt, double* mu) {
    const bool ite = (e->beta != 0.0);
    for (int s = 0; s < nt; s++) {
        const Phys<T> ph = base_args<T>(e).ph;
        dipolar_potential_t<T>(e);
        launch_dipolar_vstep<T>((cplx<T>*)e->A, (const cplx<T>*)e->B, ph, e->n, e->stream);
        kinetic_apply<T>(e, e->A, false);
        dipolar_potential_t<T>(e);
        launch_dipolar_vstep<T>((cplx<T>*)e->A, (const cplx<T>*)e->B, ph, e->n, e->stream);
        if (ite) {
            int r = normalize_t<T>(e);
            if (r) return r;
        }
        e->it = e->it + 1;
    }
    HIP_TRY(hipGetLastError());
    if (ite && mu) {
        HIP_TRY(hipMemcpyAsync(mu, e->scal + 2, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

// method 1: sum_k Vpp(k) |rho_k|^2 / N (dev 1716-1760, -k taken modulo the
// lattice; for the real density Re(rho(-k) Vpp rho(k)) = Vpp |rho(k)|^2);
// method 2: (1/2) sum rho V_dip (dev 1768-1810)
template <typename T>
int energy_dipolar_t(gpe_engine_s* e, int method, double* t, double* edip) {
    *t = e->t0 + e->dt * e->it;
    if (method == 1) {
        launch_density_c<T>((const cplx<T>*)e->A, (cplx<T>*)e->B, e->gamma, e->n, e->stream);
        PassArgs<T> a = base_args<T>(e);
        void* B = e->B;
        run_pass<T>(e, a, {2, B, B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        run_pass<T>(e, a, {1, B, B, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
        const int nb = run_pass<T>(e, a, {0, B, nullptr, -1, 0, LOAD_PLAIN, MID_VPP_SUM, STORE_NONE, 0, 1});
        finalize(e, nb, 1, FIN_STORE, kSlotPot, (double)e->n);
    } else {
        dipolar_potential_t<T>(e);
        const int nb = launch_dipolar_e2<T>((const cplx<T>*)e->A, (const cplx<T>*)e->B, e->gamma, e->n,
                                            e->partials, e->stream);
        finalize(e, nb, 1, FIN_STORE, kSlotPot, 2.0);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(edip, e->scal + kSlotPot, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

}  // namespace

extern "C" {

int gpe_engine_evolve_dipolar(gpe_engine_t e, int nt, double* chemical_potential) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (e->interactions != INT_DIPOLAR || e->batch != 1 || nt < 0) return GPE_ERR_ARG;
    if (int r = ready_to_launch(e)) return r;
    return DISPATCH(e, evolve_dipolar_t, e, nt, chemical_potential);
}

int gpe_engine_energy_dipolar(gpe_engine_t e, int method, double* t, double* edip) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (e->interactions != INT_DIPOLAR || e->batch != 1 || (method != 1 && method != 2) || !t || !edip)
        return GPE_ERR_ARG;
    if (int r = ready_to_launch(e)) return r;
    return DISPATCH(e, energy_dipolar_t, e, method, t, edip);
}

}  // extern "C"

// ====================================================================== adjoint
//
// Gradient of L = 1 - |<psi_t|psi(T)>|^2 / (|psi_t|^2 |psi(0)|^2) with respect
// to V_ext at every point for nt real-time steps (gpe_adjoint.hip holds the
// pointwise parts and the derivation).  The forward run keeps the state every
// C steps (C = ceil(sqrt(nt)) by default); the backward run recomputes one
// segment of C states at a time from its checkpoint and walks it backwards:
// memory C + nt/C states, work one extra forward run.

namespace {

int adj_alloc(void** p, size_t* have, size_t want, size_t unit) {
    if (*have >= want && *p) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    HIP_TRY(hipMalloc(p, want * unit));
    *have = want;
    return 0;
}

// psi <- K psi (adjoint: conj K) on a buffer, in place
template <typename T>
void kinetic_apply(gpe_engine_s* e, void* buf, bool adjoint) {
    PassArgs<T> a = base_args<T>(e);
    run_pass<T>(e, a, {2, buf, buf, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {1, buf, buf, -1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {0, buf, buf, -1, 1, LOAD_PLAIN, adjoint ? MID_EXPTC : MID_EXPT, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {1, buf, buf, 1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
    run_pass<T>(e, a, {2, buf, buf, 1, 0, LOAD_PLAIN, MID_NONE, STORE_PLAIN, 0, 0});
}

template <typename T>
AdjArgs adj_args(gpe_engine_s* e) {
    AdjArgs a;
    std::memset(&a, 0, sizeof(a));
    a.interactions = e->interactions;
    a.gamma = e->gamma;
    for (int i = 0; i < kMaxUserParams; i++) a.up[i] = e->up[i];
    a.cx = e->c.real();
    a.cy = e->c.imag();
    a.n = e->n;
    a.vx = e->adj_vx;
    a.phi = e->adj_phi;
    a.lam = e->adj_lam;
    a.lam0 = e->adj_lam0;
    a.vb = e->adj_vb;
    a.grad = e->adj_grad;
    return a;
}

template <typename T>
int adjoint_forward_t(gpe_engine_s* e, int nt, const cuCplx* tgt, int ckpt, double* loss) {
    const size_t n = e->n, cs = sizeof(cplx<T>);
    const int C = ckpt > 0 ? ckpt : (int)std::ceil(std::sqrt((double)nt));
    const int nck = (nt + C - 1) / C;
    size_t have;
    e->adj_ready = false;
    int r = adj_alloc(&e->adj_ck, &e->adj_ck_len, (size_t)nck * n, cs);
    if (!r) r = adj_alloc(&e->adj_seg, &e->adj_seg_len, (size_t)(C + 1) * n, cs);
    for (void** p : {&e->adj_lam, &e->adj_phi, &e->adj_lam0, &e->adj_tgt, &e->adj_fin}) {
        have = *p ? n : 0;
        if (!r) r = adj_alloc(p, &have, n, cs);
    }
    for (double** p : {&e->adj_vb, &e->adj_grad, &e->adj_vx}) {
        have = *p ? n : 0;
        if (!r) r = adj_alloc((void**)p, &have, n, sizeof(double));
    }
    if (r) return r;
    // target in the working precision, V_ext as the forward kernels see it
    {
        std::vector<cplx<T>> t(n);
        for (size_t i = 0; i < n; i++) t[i] = mk((T)tgt[i].x, (T)tgt[i].y);
        HIP_TRY(hipMemcpyAsync(e->adj_tgt, t.data(), n * cs, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    launch_adjoint_vfill<T>(base_args<T>(e).ph, e->adj_vx, n, e->stream);
    e->adj_it0 = e->it;
    for (int k = 0; k < nt; k++) {
        if (k % C == 0)
            HIP_TRY(hipMemcpyAsync((char*)e->adj_ck + (size_t)(k / C) * n * cs, e->A, n * cs, hipMemcpyDeviceToDevice,
                                   e->stream));
        r = evolve_t<T>(e, 1, nullptr, 0, false);
        if (r) return r;
    }
    const int nb = launch_adjoint_overlap<T>(e->A, e->adj_tgt, e->adj_ck, n, e->partials, e->stream);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(4 * (size_t)nb);
    HIP_TRY(hipMemcpyAsync(h.data(), e->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    double s4[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; k++)
        for (int b = 0; b < nb; b++) s4[k] += h[(size_t)k * nb + b];
    const double sc = 1.0 / (s4[2] * s4[3]);
    *loss = 1.0 - (s4[0] * s4[0] + s4[1] * s4[1]) * sc;
    e->adj_fr = -sc * s4[0];
    e->adj_fi = -sc * s4[1];
    e->adj_nt = nt;
    e->adj_C = C;
    e->adj_gen = e->gen;
    e->adj_ready = true;
    return 0;
}

template <typename T>
int adjoint_backward_t(gpe_engine_s* e, double* grad, bool device_out = false) {
    const size_t n = e->n, cs = sizeof(cplx<T>);
    const int nt = e->adj_nt, C = e->adj_C;
    const int nck = (nt + C - 1) / C;
    const unsigned it_end = e->it;
    HIP_TRY(hipMemcpyAsync(e->adj_fin, e->A, n * cs, hipMemcpyDeviceToDevice, e->stream));
    launch_adjoint_lam_init<T>(e->adj_lam, e->adj_tgt, e->adj_fr, e->adj_fi, n, e->stream);
    HIP_TRY(hipMemsetAsync(e->adj_grad, 0, n * sizeof(double), e->stream));
    AdjArgs a = adj_args<T>(e);
    for (int sg = nck - 1; sg >= 0; sg--) {
        const int k0 = sg * C, k1 = std::min(nt, k0 + C);
        char* seg = (char*)e->adj_seg;
        HIP_TRY(hipMemcpyAsync(e->A, (char*)e->adj_ck + (size_t)sg * n * cs, n * cs, hipMemcpyDeviceToDevice,
                               e->stream));
        HIP_TRY(hipMemcpyAsync(seg, e->A, n * cs, hipMemcpyDeviceToDevice, e->stream));
        e->it = e->adj_it0 + (unsigned)k0;
        for (int j = 1; j <= k1 - k0; j++) {
            int r = evolve_t<T>(e, 1, nullptr, 0, false);
            if (r) return r;
            HIP_TRY(hipMemcpyAsync(seg + (size_t)j * n * cs, e->A, n * cs, hipMemcpyDeviceToDevice, e->stream));
        }
        for (int k = k1 - 1; k >= k0; k--) {
            a.psi0 = seg + (size_t)(k - k0) * n * cs;
            a.psi1 = seg + (size_t)(k - k0 + 1) * n * cs;
            launch_adjoint_step<T>(a, 1, e->stream);
            kinetic_apply<T>(e, e->adj_phi, false);
            launch_adjoint_step<T>(a, 2, e->stream);
            kinetic_apply<T>(e, e->adj_lam, true);
            launch_adjoint_step<T>(a, 3, e->stream);
            kinetic_apply<T>(e, e->adj_lam, true);
            launch_adjoint_step<T>(a, 4, e->stream);
        }
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(e->A, e->adj_fin, n * cs, hipMemcpyDeviceToDevice, e->stream));
    e->it = it_end;
    HIP_TRY(hipMemcpyAsync(grad, e->adj_grad, n * sizeof(double),
                           device_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

bool adjoint_supported(const gpe_engine_s* e) {
    return e->beta == 0.0 && e->qf == 0.0 && e->batch == 1 && e->vext != VEXT_USER &&
           (e->interactions == INT_UFG || e->interactions == INT_CONTACT || e->interactions == INT_DIPOLAR);
}

}  // namespace

extern "C" {

int gpe_engine_adjoint_forward(gpe_engine_t e, int nt, const cuCplx* psi_target, int ckpt, double* loss) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (nt < 1 || !psi_target || !loss || ckpt < 0 || !adjoint_supported(e)) return GPE_ERR_ARG;
    if (int r = ready_to_launch(e)) return r;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return DISPATCH(e, adjoint_forward_t, e, nt, psi_target, ckpt, loss);
}

int gpe_engine_adjoint_backward(gpe_engine_t e, double* grad_vext) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!grad_vext || !e->adj_ready || !adjoint_supported(e)) return GPE_ERR_ARG;
    // the backward run recomputes the forward trajectory with the engine's
    // current parameters: they must be those of the forward run
    if (e->adj_gen != e->gen) return GPE_ERR_ARG;
    return DISPATCH(e, adjoint_backward_t, e, grad_vext);
}

int gpe_engine_adjoint_backward_device(gpe_engine_t e, double* d_grad_vext) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (!d_grad_vext || !e->adj_ready || !adjoint_supported(e)) return GPE_ERR_ARG;
    if (e->adj_gen != e->gen) return GPE_ERR_ARG;
    return DISPATCH(e, adjoint_backward_t, e, d_grad_vext, true);
}

}  // extern "C"

// ===================================================================== ensembles
//
// An ensemble is a set of independent lattices (truncated-Wigner noise
// realisations, parameter sweeps) advanced together in real time.  Lattices of
// one shape form a group: one engine whose state arrays hold the group's
// lattices one after another, so every fused pass of a step covers the whole
// group in one launch (pencil addressing runs across lattice boundaries).  A
// step of the ensemble issues each group's passes on the group's own stream,
// forked from and joined back into the ensemble stream, and is captured once
// as a HIP graph: open, middle and last step graphs, replayed nt - 2 times for
// the middle.  Kernel arguments are frozen in the graphs, so the graphs are
// rebuilt after every parameter change; the user hooks read the time index
// from a device counter advanced after every step, not from the arguments.
// Groups on the fused whole-step kernel run outside the graphs, on their own
// streams, for all nt steps of a call.

struct gpe_ensemble_s {
    std::vector<gpe_engine_s*> groups;
    std::vector<int> inst_group;
    std::vector<size_t> inst_slot;
    int precision = GPE_PRECISION_C64;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};  // open, middle step, last step
    bool graphs_ok = false;
    bool use_graph = true;
};

namespace {

void ens_drop_graphs(gpe_ensemble_s* s) {
    for (auto& g : s->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    s->graphs_ok = false;
}

void ens_free(gpe_ensemble_s* s) {
    ens_drop_graphs(s);
    for (auto* g : s->groups) {
        if (g->stream) (void)hipStreamSynchronize(g->stream);
        free_all(g);
        delete g;
    }
    for (auto ev : s->join) (void)hipEventDestroy(ev);
    if (s->fork) (void)hipEventDestroy(s->fork);
    if (s->main) (void)hipStreamDestroy(s->main);
    delete s;
}

// what: 0 open pass, 1 a middle step, 2 the last step
template <typename T>
void ens_issue(gpe_ensemble_s* s, int what) {
    (void)hipEventRecord(s->fork, s->main);
    StepPlan sp;
    for (size_t i = 0; i < s->groups.size(); i++) {
        gpe_engine_s* g = s->groups[i];
        if (fused_selected<T>(g, 0)) continue;  // whole steps in one kernel, launched by ens_evolve_t
        (void)hipStreamWaitEvent(g->stream, s->fork, 0);
        if (what == 0) issue_open<T>(g, sp);
        else issue_step<T>(g, sp, what == 2);
        (void)hipEventRecord(s->join[i], g->stream);
        (void)hipStreamWaitEvent(s->main, s->join[i], 0);
    }
}

template <typename T>
int ens_build_graphs(gpe_ensemble_s* s) {
    ens_drop_graphs(s);
    for (int w = 0; w < 3; w++) {
        HIP_TRY(hipStreamBeginCapture(s->main, hipStreamCaptureModeThreadLocal));
        ens_issue<T>(s, w);
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamEndCapture(s->main, &g));
        const hipError_t err = hipGraphInstantiate(&s->graph[w], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_TRY(err);
    }
    s->graphs_ok = true;
    return 0;
}

template <typename T>
int ens_evolve_t(gpe_ensemble_s* s, int nt) {
    if (nt <= 0) return 0;
    for (auto* g : s->groups) {
        if (g->beta != 0.0) return GPE_ERR_ARG;  // real time only: one norm per lattice is not kept
        if (int r = ready_to_launch(g)) return r;
    }
    for (auto* g : s->groups) HIP_TRY(hipStreamSynchronize(g->stream));
    for (auto* g : s->groups)
        if (g->uses_hooks) launch_set_counter(g->itdev, g->it, s->main);
    // groups on the fused whole-step kernel: nt steps on the group's stream,
    // beside the passes of the other groups
    (void)hipEventRecord(s->fork, s->main);
    for (size_t i = 0; i < s->groups.size(); i++) {
        gpe_engine_s* g = s->groups[i];
        if (!fused_selected<T>(g, 0)) continue;
        HIP_TRY(hipStreamWaitEvent(g->stream, s->fork, 0));
        if (int r = fused_run<T>(g, nt, g->stream)) return r;
        HIP_TRY(hipEventRecord(s->join[i], g->stream));
    }
    if (s->use_graph) {
        if (!s->graphs_ok) {
            int r = ens_build_graphs<T>(s);
            if (r) return r;
        }
        HIP_TRY(hipGraphLaunch(s->graph[0], s->main));
        for (int i = 0; i < nt - 1; i++) HIP_TRY(hipGraphLaunch(s->graph[1], s->main));
        HIP_TRY(hipGraphLaunch(s->graph[2], s->main));
    } else {
        ens_issue<T>(s, 0);
        for (int i = 0; i < nt - 1; i++) ens_issue<T>(s, 1);
        ens_issue<T>(s, 2);
    }
    for (size_t i = 0; i < s->groups.size(); i++)
        if (fused_selected<T>(s->groups[i], 0)) HIP_TRY(hipStreamWaitEvent(s->main, s->join[i], 0));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->main));
    for (auto* g : s->groups) g->it += (unsigned)nt;
    return 0;
}

// Times `steps` real-time steps of one engine for each launcher variant
// (bit 0: z passes in LDS tiles instead of wave registers, bit 1: half-width
// strided tiles, bit 2: pipelined strided passes) and keeps the fastest; bit 3
// (the unsplit kinetic plan) is kept as set.  Then the alternative radix
// schedule of the strided turn-arounds (bit 6) and the y-z plane plan (bit 5)
// are tried on top of the fastest, and where the fused whole-step kernel
// applies, bit 4 is added when it beats the rest.  The state is saved and
// restored.
template <typename T>
int autotune_t(gpe_engine_s* e, int steps, int* chosen) {
    if (steps < 1) steps = 1;
    const size_t bytes = e->n * e->batch * sizeof(cplx<T>);
    void* save = nullptr;
    HIP_TRY(hipMalloc(&save, bytes));
    const unsigned it0 = e->it;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    int best = 0;
    float best_ms = 0.f;
    int err = 0;
    (void)hipMemcpyAsync(save, e->A, bytes, hipMemcpyDeviceToDevice, e->stream);
    const double beta = e->beta;
    const int plan = e->variant & 8;
    for (int v = 0; v < 8 && !err; v++) {
        if ((v & 4) && e->precision != GPE_PRECISION_C64) continue;  // pipelined passes: complex64 only
        e->variant = v | plan;
        StepPlan sp;
        // one untimed step, then the timed ones (real-time schedule in both time modes)
        e->beta = 0.0;
        issue_open<T>(e, sp);
        issue_step<T>(e, sp, true);
        (void)hipEventRecord(t0, e->stream);
        issue_open<T>(e, sp);
        for (int i = 0; i < steps; i++) issue_step<T>(e, sp, i == steps - 1);
        (void)hipEventRecord(t1, e->stream);
        e->beta = beta;
        hipError_t h = hipEventSynchronize(t1);
        if (h != hipSuccess) { err = (int)h; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        if (v == 0 || ms < best_ms) { best = v | plan; best_ms = ms; }
    }
    // the alternative radix schedule of the strided turn-arounds (variant bit 6)
    const bool c64 = e->precision == GPE_PRECISION_C64;
    if (!err && !(best & 8) && (has_alt_plan(e->nx, c64) || has_alt_plan(e->ny, c64))) {
        const int cand = best | kVariantAltPlan;
        e->variant = cand;
        StepPlan sp;
        e->beta = 0.0;
        issue_open<T>(e, sp);
        issue_step<T>(e, sp, true);
        (void)hipEventRecord(t0, e->stream);
        issue_open<T>(e, sp);
        for (int i = 0; i < steps; i++) issue_step<T>(e, sp, i == steps - 1);
        (void)hipEventRecord(t1, e->stream);
        e->beta = beta;
        const hipError_t h = hipEventSynchronize(t1);
        float ms = 0.f;
        if (h != hipSuccess) err = (int)h;
        else if (hipEventElapsedTime(&ms, t0, t1) == hipSuccess && ms < best_ms) { best = cand; best_ms = ms; }
    }
    // the plane plan where it applies (its x passes keep the best bits)
    if (!err && plane_eligible(e, 0)) {
        const int cand = best | kVariantPlane;
        e->variant = cand;
        StepPlan sp;
        e->beta = 0.0;
        issue_open<T>(e, sp);
        issue_step<T>(e, sp, true);
        (void)hipEventRecord(t0, e->stream);
        issue_open<T>(e, sp);
        for (int i = 0; i < steps; i++) issue_step<T>(e, sp, i == steps - 1);
        (void)hipEventRecord(t1, e->stream);
        e->beta = beta;
        const hipError_t h = hipEventSynchronize(t1);
        float ms = 0.f;
        if (h != hipSuccess) err = (int)h;
        else if (hipEventElapsedTime(&ms, t0, t1) == hipSuccess && ms < best_ms) { best = cand; best_ms = ms; }
    }
    e->beta = beta;
    // the fused whole-step kernel where it applies, keeping the best passes
    // for the runs it does not cover
    if (!err && fused_eligible<T>(e, 0)) {
        e->variant = best | kVariantFused;
        err = fused_run<T>(e, 1, e->stream);
        (void)hipEventRecord(t0, e->stream);
        if (!err) err = fused_run<T>(e, steps, e->stream);
        (void)hipEventRecord(t1, e->stream);
        if (!err) {
            const hipError_t h = hipEventSynchronize(t1);
            float ms = 0.f;
            if (h != hipSuccess) err = (int)h;
            else if (hipEventElapsedTime(&ms, t0, t1) == hipSuccess && ms < best_ms) best |= kVariantFused;
        }
    }
    (void)hipMemcpyAsync(e->A, save, bytes, hipMemcpyDeviceToDevice, e->stream);
    (void)hipStreamSynchronize(e->stream);
    (void)hipFree(save);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    e->it = it0;
    e->variant = best;
    if (chosen) *chosen = best;
    return err;
}

}  // namespace

extern "C" {

int gpe_engine_autotune(gpe_engine_t e, int steps, int* variant) {
    if (!e) return GPE_ERR_NO_ENGINE;
    if (int r = ready_to_launch(e)) return r;
    e->gen++;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return DISPATCH(e, autotune_t, e, steps, variant);
}

int gpe_engine_set_graph(gpe_engine_t e, int on) {
    if (!e) return GPE_ERR_NO_ENGINE;
    e->use_graph = (on != 0);
    if (!e->use_graph) drop_graphs(e);
    return 0;
}

int gpe_engine_set_variant(gpe_engine_t e, int variant) {
    if (!e) return GPE_ERR_NO_ENGINE;
    e->gen++;
    if (variant < 0 || variant > 127) return GPE_ERR_ARG;
    e->variant = variant;
    return 0;
}

int gpe_ensemble_create(gpe_ensemble_t* out, int ninst, const int* dims, int interactions, int vext, int gpe_for,
                        int precision, double alpha, double beta, double dt, double npart) {
    if (!out || ninst < 1 || !dims) return GPE_ERR_ARG;
    *out = nullptr;
    // group lattices by shape, in order of first appearance
    std::vector<std::array<int, 3>> shapes;
    std::vector<size_t> counts;
    std::vector<int> grp(ninst);
    std::vector<size_t> slot(ninst);
    for (int i = 0; i < ninst; i++) {
        std::array<int, 3> d = {dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]};
        size_t g = 0;
        while (g < shapes.size() && shapes[g] != d) g++;
        if (g == shapes.size()) {
            shapes.push_back(d);
            counts.push_back(0);
        }
        grp[i] = (int)g;
        slot[i] = counts[g]++;
    }
    gpe_ensemble_s* s = new gpe_ensemble_s();
    s->precision = precision;
    s->inst_group = grp;
    s->inst_slot = slot;
    for (size_t g = 0; g < shapes.size(); g++) {
        gpe_config_t cfg{shapes[g][0], shapes[g][1], shapes[g][2], interactions, vext, gpe_for, precision};
        gpe_engine_s* e = nullptr;
        const int r = create_impl(&e, &cfg, alpha, beta, dt, npart, counts[g]);
        if (r) {
            ens_free(s);
            return r;
        }
        s->groups.push_back(e);
    }
    hipError_t err = hipStreamCreateWithFlags(&s->main, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&s->fork, hipEventDisableTiming);
    for (size_t g = 0; g < shapes.size() && err == hipSuccess; g++) {
        hipEvent_t ev;
        err = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (err == hipSuccess) s->join.push_back(ev);
    }
    if (err != hipSuccess) {
        ens_free(s);
        return (int)err;
    }
    *out = s;
    return 0;
}

int gpe_ensemble_destroy(gpe_ensemble_t s) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (s->main) (void)hipStreamSynchronize(s->main);
    ens_free(s);
    return 0;
}

int gpe_ensemble_set_user_params(gpe_ensemble_t s, int size, const double* params) {
    if (!s) return GPE_ERR_NO_ENGINE;
    ens_drop_graphs(s);
    for (auto* g : s->groups) {
        const int r = gpe_engine_set_user_params(g, size, params);
        if (r) return r;
    }
    return 0;
}

int gpe_ensemble_set_quantum_friction_coeff(gpe_ensemble_t s, double qf) {
    if (!s) return GPE_ERR_NO_ENGINE;
    ens_drop_graphs(s);
    for (auto* g : s->groups) {
        const int r = gpe_engine_set_quantum_friction_coeff(g, qf);
        if (r) return r;
    }
    return 0;
}

int gpe_ensemble_set_instance_params(gpe_ensemble_t s, int inst, int size, const double* params) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (inst < 0 || inst >= (int)s->inst_group.size()) return GPE_ERR_ARG;
    if (size > MAX_USER_PARAMS || size > kMaxUserParams) return GPE_ERR_TOO_MANY_PARAMS;
    if (size < 0 || (size > 0 && !params)) return GPE_ERR_ARG;
    gpe_engine_s* g = s->groups[s->inst_group[inst]];
    // the built-in physics reads one parameter set per group; the user hooks
    // read d_user_param per lattice
    if (!g->uses_hooks) return GPE_ERR_ARG;
    double up[kMaxUserParams];
    for (int i = 0; i < kMaxUserParams; i++) up[i] = (i < size) ? params[i] : 0.0;
    ens_drop_graphs(s);
    g->gen++;
    return upload_user_params(g, (long)s->inst_slot[inst], up);
}

int gpe_ensemble_set_external_potential(gpe_ensemble_t s, int inst, const double* v) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (inst < 0 || inst >= (int)s->inst_group.size() || !v) return GPE_ERR_ARG;
    gpe_engine_s* g = s->groups[s->inst_group[inst]];
    if (g->vext != VEXT_ARRAY) return GPE_ERR_ARG;
    ens_drop_graphs(s);
    return set_varr_slots(g, (long)s->inst_slot[inst], v);
}

int gpe_ensemble_set_psi(gpe_ensemble_t s, int inst, const cuCplx* psi) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (inst < 0 || inst >= (int)s->inst_group.size() || !psi) return GPE_ERR_ARG;
    gpe_engine_s* g = s->groups[s->inst_group[inst]];
    g->t0 = 0.0;
    g->it = 0;
    return DISPATCH(g, set_psi_t, g, psi, s->inst_slot[inst]);
}

int gpe_ensemble_get_psi(gpe_ensemble_t s, int inst, double* t, cuCplx* psi) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (inst < 0 || inst >= (int)s->inst_group.size() || !psi) return GPE_ERR_ARG;
    gpe_engine_s* g = s->groups[s->inst_group[inst]];
    if (t) *t = g->t0 + g->dt * g->it;
    return DISPATCH(g, get_psi_t, g, psi, s->inst_slot[inst]);
}

int gpe_ensemble_evolve(gpe_ensemble_t s, int nt) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (nt < 0) return GPE_ERR_ARG;
    return s->precision == GPE_PRECISION_C64 ? ens_evolve_t<float>(s, nt) : ens_evolve_t<double>(s, nt);
}

int gpe_ensemble_set_graph(gpe_ensemble_t s, int on) {
    if (!s) return GPE_ERR_NO_ENGINE;
    s->use_graph = (on != 0);
    return 0;
}

int gpe_ensemble_autotune(gpe_ensemble_t s, int steps) {
    if (!s) return GPE_ERR_NO_ENGINE;
    for (auto* g : s->groups)
        if (int r = ready_to_launch(g)) return r;
    ens_drop_graphs(s);
    for (auto* g : s->groups) {
        const int r = gpe_engine_autotune(g, steps, nullptr);
        if (r) return r;
    }
    return 0;
}

int gpe_ensemble_set_variant(gpe_ensemble_t s, int group, int variant) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (group < 0 || group >= (int)s->groups.size()) return GPE_ERR_ARG;
    ens_drop_graphs(s);
    return gpe_engine_set_variant(s->groups[group], variant);
}

int gpe_ensemble_groups(gpe_ensemble_t s, int* ngroups) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (!ngroups) return GPE_ERR_ARG;
    *ngroups = (int)s->groups.size();
    return 0;
}

int gpe_ensemble_group_info(gpe_ensemble_t s, int group, int* nx, int* ny, int* nz, int* count, int* variant) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (group < 0 || group >= (int)s->groups.size()) return GPE_ERR_ARG;
    const gpe_engine_s* g = s->groups[group];
    if (nx) *nx = g->nx;
    if (ny) *ny = g->ny;
    if (nz) *nz = g->nz;
    if (count) *count = (int)g->batch;
    if (variant) *variant = g->variant;
    return 0;
}

int gpe_ensemble_instance(gpe_ensemble_t s, int inst, int* group, int* slot) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (inst < 0 || inst >= (int)s->inst_group.size()) return GPE_ERR_ARG;
    if (group) *group = s->inst_group[inst];
    if (slot) *slot = (int)s->inst_slot[inst];
    return 0;
}

int gpe_ensemble_mean_density(gpe_ensemble_t s, int group, double* out) {
    if (!s) return GPE_ERR_NO_ENGINE;
    if (group < 0 || group >= (int)s->groups.size() || !out) return GPE_ERR_ARG;
    gpe_engine_s* g = s->groups[group];
    int r = ensure_rout(g);
    if (r) return r;
    if (g->precision == GPE_PRECISION_C64)
        launch_mean_density<float>((const cplx<float>*)g->A, (unsigned)g->n, (unsigned)g->batch, g->gamma, g->rout,
                                   g->stream);
    else
        launch_mean_density<double>((const cplx<double>*)g->A, (unsigned)g->n, (unsigned)g->batch, g->gamma, g->rout,
                                    g->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, g->rout, g->n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return 0;
}

}  // extern "C"
