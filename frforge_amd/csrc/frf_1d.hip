// 1D Euler path (the Sod tube configuration) on the GPU: Euler1D residual with Rusanov
// interface fluxes, Persson sensor with BR0 or element-local artificial viscosity, CFL step and
// SSP-RK3, in FP64 with the reference's operation order and no fused multiply-adds
// (Residual.jl:24-214, PerssonAV.jl:103-352, SharedAV.jl:9-42, SSP_RK3.jl:8-118, Euler.jl).
//
// The 1D problems of the reference are small (tens to hundreds of elements), so a whole run is
// ONE launch of one workgroup: threads sweep the elements (and the Nel+1 interfaces) in strided
// loops, phases are separated by workgroup barriers, the step size comes from a workgroup
// reduction, and the time loop itself runs on the device.  Phase scratch (traces, interface
// fluxes, viscosities) lives in device memory owned by the context.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/frforge_hip.h"
#include "frf_kernels.h"

namespace frf {
namespace d1 {

constexpr int kNeq1 = 3;
constexpr int kThreads = 256;
constexpr double kEps = 2.220446049250313080847e-16;

struct Ops1 {
    double D[kMaxNp * kMaxNp];   // D[j + Np*k] = D[j,k] (column-major, as the reference stores it)
    double lL[kMaxNp], lR[kMaxNp], gL[kMaxNp], gR[kMaxNp];
    double LU[kMaxNp * kMaxNp];  // LU of the Legendre Vandermonde with partial pivoting
    int piv[kMaxNp];
};

struct Ctl1 {
    double t, t_final, cfl, fixed_dt;
    int64_t n_steps, max_steps;
    int32_t status, fail_stage, mode, pad;   // mode 0 residual, 1 one step of dt = fixed_dt, 2 run
};

struct Par1 {
    double* u;                 // state (in / out)
    double *u0, *u1, *du;      // stage buffers
    double *uL, *uR, *fL, *fR, *sig, *eps, *g, *gL, *gR, *ghL, *ghR;
    const double* dx;          // element widths
    const double* J;           // dx / 2
    const Ops1* op;
    Ctl1* ctl;
    int np, nel, periodic, capt;
    int dir_l, dir_r;          // Dirichlet ends (constant ghost states ud_l, ud_r)
    double ud_l[3], ud_r[3];
    double gamma, kappa, s0, eps_floor, c_av;
};

__device__ __forceinline__ double jmax(double a, double b) {
    if (a != a || b != b) return a + b;   // NaN propagates, as Julia's max
    return a > b ? a : b;
}
__device__ __forceinline__ double jmin(double a, double b) {
    if (a != a || b != b) return a + b;
    return a < b ? a : b;
}

__device__ double pres(double g, double rho, double ru, double E) {
#pragma clang fp contract(off)
    const double rs = jmax(rho, kEps);
    const double u = ru / rs;
    return (g - 1.0) * (E - 0.5 * rs * u * u);
}
__device__ double snd(double g, double rho, double ru, double E) {
#pragma clang fp contract(off)
    const double r = jmax(rho, kEps);
    const double p = jmax(pres(g, rho, ru, E), kEps);
    return sqrt(g * p / r);
}
__device__ double vel(double rho, double ru) {
#pragma clang fp contract(off)
    return ru / jmax(rho, kEps);
}
__device__ void pflux(double g, const double U[3], double F[3]) {
#pragma clang fp contract(off)
    const double u = U[1] / U[0];
    const double p = pres(g, U[0], U[1], U[2]);
    F[0] = U[1];
    F[1] = U[1] * u + p;
    F[2] = (U[2] + p) * u;
}
__device__ void rusanov(double g, const double uL[3], const double uR[3], double out[3]) {
#pragma clang fp contract(off)
    double fL[3], fR[3];
    pflux(g, uL, fL);
    pflux(g, uR, fR);
    const double lam = jmax(fabs(vel(uL[0], uL[1])) + snd(g, uL[0], uL[1], uL[2]),
                            fabs(vel(uR[0], uR[1])) + snd(g, uR[0], uR[1], uR[2]));
    for (int c = 0; c < 3; ++c) out[c] = 0.5 * (fL[c] + fR[c]) - 0.5 * lam * (uR[c] - uL[c]);
}

#define UU(a, j, e, c) (a)[(j) + Np * ((e) + Nel * (c))]
#define EE(a, e, c) (a)[(e) + Nel * (c)]

// residual! (Residual.jl:161-214) of u into P.du; every thread of the workgroup takes part
__device__ void residual(const Par1& P, const double* __restrict__ u) {
#pragma clang fp contract(off)
    const int Np = P.np, Nel = P.nel, tid = threadIdx.x;
    const Ops1& op = *P.op;
    const double g = P.gamma;
    for (int e = tid; e < Nel; e += kThreads)   // traces (Interfaces.jl:49-72)
        for (int c = 0; c < 3; ++c) {
            double sL = 0.0, sR = 0.0;
            for (int j = 0; j < Np; ++j) {
                sL += op.lL[j] * UU(u, j, e, c);
                sR += op.lR[j] * UU(u, j, e, c);
            }
            EE(P.uL, e, c) = sL;
            EE(P.uR, e, c) = sR;
        }
    __syncthreads();
    // interface i sits left of element i (i = 0..Nel): one Rusanov flux each (Residual.jl:24-115)
    for (int i = tid; i <= Nel; i += kThreads) {
        double um[3], up[3], f[3];
        if (i > 0 && i < Nel) {
            for (int c = 0; c < 3; ++c) { um[c] = EE(P.uR, i - 1, c); up[c] = EE(P.uL, i, c); }
            rusanov(g, um, up, f);
            for (int c = 0; c < 3; ++c) { EE(P.fR, i - 1, c) = f[c]; EE(P.fL, i, c) = f[c]; }
        } else if (i == 0) {   // Residual.jl:63-90: periodic, transmissive or Dirichlet ghost
            for (int c = 0; c < 3; ++c) {
                up[c] = EE(P.uL, 0, c);
                um[c] = P.periodic ? EE(P.uR, Nel - 1, c) : (P.dir_l ? P.ud_l[c] : EE(P.uL, 0, c));
            }
            rusanov(g, um, up, f);
            for (int c = 0; c < 3; ++c) {
                EE(P.fL, 0, c) = f[c];
                if (P.periodic) EE(P.fR, Nel - 1, c) = f[c];
            }
        } else if (!P.periodic) {   // i == Nel
            for (int c = 0; c < 3; ++c) {
                um[c] = EE(P.uR, Nel - 1, c);
                up[c] = P.dir_r ? P.ud_r[c] : EE(P.uR, Nel - 1, c);
            }
            rusanov(g, um, up, f);
            for (int c = 0; c < 3; ++c) EE(P.fR, Nel - 1, c) = f[c];
        }
    }
    __syncthreads();
    for (int e = tid; e < Nel; e += kThreads) {   // volume + corrections
        double F[kMaxNp][3], Uj[3], fLc[3], fRc[3];
        for (int j = 0; j < Np; ++j) {
            for (int c = 0; c < 3; ++c) Uj[c] = UU(u, j, e, c);
            pflux(g, Uj, F[j]);
        }
        for (int c = 0; c < 3; ++c) {
            double sL = 0.0, sR = 0.0;
            for (int j = 0; j < Np; ++j) sL += op.lL[j] * F[j][c];
            for (int j = 0; j < Np; ++j) sR += op.lR[j] * F[j][c];
            fLc[c] = sL;
            fRc[c] = sR;
        }
        const double Je = P.J[e];
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < Np; ++j) {
                double vol = 0.0;
                for (int k = 0; k < Np; ++k) vol += op.D[j + Np * k] * F[k][c];
                const double corr = (EE(P.fL, e, c) - fLc[c]) * op.gL[j] + (EE(P.fR, e, c) - fRc[c]) * op.gR[j];
                UU(P.du, j, e, c) = -(vol + corr) / Je;
            }
    }
    if (P.capt == 0) {
        __syncthreads();
        return;
    }
    // Persson sensor (PerssonAV.jl:116-135) and element viscosities (PerssonAV.jl:155-176)
    const int pp = P.np - 1 > 1 ? P.np - 1 : 1;
    for (int e = tid; e < Nel; e += kThreads) {
        double b[kMaxNp];
        for (int j = 0; j < Np; ++j) b[j] = UU(u, j, e, 0);
        for (int k = 0; k < Np; ++k) {   // V x = b by the LU factors: P L U x = b
            const int q = op.piv[k];
            if (q != k) { const double tt = b[k]; b[k] = b[q]; b[q] = tt; }
        }
        for (int i = 0; i < Np; ++i)
            for (int k = 0; k < i; ++k) b[i] -= op.LU[i + Np * k] * b[k];
        for (int i = Np - 1; i >= 0; --i) {
            for (int k = i + 1; k < Np; ++k) b[i] -= op.LU[i + Np * k] * b[k];
            b[i] /= op.LU[i + Np * i];
        }
        const double eh = b[Np - 1] * b[Np - 1];
        double et = 0.0;
        for (int j = 0; j < Np; ++j) et += b[j] * b[j];
        et += P.eps_floor;
        const double se = log10(eh / et + P.eps_floor);
        const double sg = 1.0 / (1.0 + exp(-P.kappa * (se - P.s0)));
        P.sig[e] = sg;
        double ev = 0.0;
        if (sg > 0.0) {
            double m = 0.0;
            for (int j = 0; j < Np; ++j) {
                const double r = UU(u, j, e, 0), ru = UU(u, j, e, 1), E = UU(u, j, e, 2);
                m = jmax(m, fabs(vel(r, ru)) + snd(g, r, ru, E));
            }
            m = jmax(m, kEps);
            ev = P.c_av * sg * (P.dx[e] / double(pp)) * m;
        }
        P.eps[e] = ev;
    }
    if (P.capt == 2) {   // apply_dissipation_local_DD! (PerssonAV.jl:194-227)
        for (int e = tid; e < Nel; e += kThreads) {
            const double ee = P.eps[e];
            if (ee <= 0.0) continue;
            const double Jj = P.J[e];
            const double scale = ee / (Jj * Jj);
            for (int c = 0; c < 3; ++c) {
                double v[kMaxNp];
                for (int j = 0; j < Np; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < Np; ++k) s += op.D[j + Np * k] * UU(u, k, e, c);
                    v[j] = s;
                }
                for (int j = 0; j < Np; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < Np; ++k) s += op.D[j + Np * k] * v[k];
                    UU(P.du, j, e, c) += scale * s;
                }
            }
        }
        __syncthreads();
        return;
    }
    // apply_dissipation_br0! (PerssonAV.jl:236-340)
    for (int e = tid; e < Nel; e += kThreads) {
        const double Jj = P.J[e], ee = P.eps[e];
        for (int c = 0; c < 3; ++c) {
            for (int j = 0; j < Np; ++j) {
                double dq = 0.0;
                for (int k = 0; k < Np; ++k) dq += op.D[j + Np * k] * UU(u, k, e, c);
                UU(P.g, j, e, c) = ee * dq / Jj;
            }
            double suL = 0.0, suR = 0.0, sgL = 0.0, sgR = 0.0;
            for (int j = 0; j < Np; ++j) {
                suL += op.lL[j] * UU(u, j, e, c);
                suR += op.lR[j] * UU(u, j, e, c);
                sgL += op.lL[j] * UU(P.g, j, e, c);
                sgR += op.lR[j] * UU(P.g, j, e, c);
            }
            EE(P.uL, e, c) = suL;
            EE(P.uR, e, c) = suR;
            EE(P.gL, e, c) = sgL;
            EE(P.gR, e, c) = sgR;
        }
    }
    __syncthreads();
    const int p = P.np - 1;
    for (int i = tid; i <= Nel; i += kThreads) {
        const int el = (i == 0) ? Nel - 1 : i - 1, er = (i == Nel) ? 0 : i;
        const bool interior = i > 0 && i < Nel;
        if (interior || (P.periodic && i == 0)) {
            for (int c = 0; c < 3; ++c) {
                const double tau = 2.0 * (0.5 * (P.eps[el] + P.eps[er])) * double(p + 1) /
                                   jmax(0.5 * (P.dx[el] + P.dx[er]), kEps);
                const double gh = 0.5 * (EE(P.gR, el, c) + EE(P.gL, er, c)) - tau * (EE(P.uL, er, c) - EE(P.uR, el, c));
                EE(P.ghR, el, c) = gh;
                EE(P.ghL, er, c) = gh;
            }
        } else if (!P.periodic) {
            for (int c = 0; c < 3; ++c) {
                if (i == 0) EE(P.ghL, 0, c) = EE(P.gL, 0, c);
                if (i == Nel) EE(P.ghR, Nel - 1, c) = EE(P.gR, Nel - 1, c);
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < Nel; e += kThreads) {
        const double Jj = P.J[e];
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < Np; ++j) {
                double vol = 0.0;
                for (int k = 0; k < Np; ++k) vol += op.D[j + Np * k] * UU(P.g, k, e, c);
                const double corr = (EE(P.ghL, e, c) - EE(P.gL, e, c)) * op.gL[j] +
                                    (EE(P.ghR, e, c) - EE(P.gR, e, c)) * op.gR[j];
                UU(P.du, j, e, c) += (vol + corr) / Jj;
            }
    }
    __syncthreads();
}

__device__ bool any_nonfinite(const double* a, int n) {
    int bad = 0;
    for (int k = threadIdx.x; k < n; k += kThreads) bad |= !isfinite(a[k]);
    return __syncthreads_or(bad) != 0;
}

// compute_dt 1D (SSP_RK3.jl:8-18): block maximum of |u| + c, then the minimum over elements
__device__ double compute_dt(const Par1& P, const double* u, double cfl, double* red) {
#pragma clang fp contract(off)
    const int Np = P.np, Nel = P.nel;
    double m = 0.0;
    for (int e = threadIdx.x; e < Nel; e += kThreads)
        for (int j = 0; j < Np; ++j) {
            const double r = UU(u, j, e, 0), ru = UU(u, j, e, 1), E = UU(u, j, e, 2);
            m = jmax(m, fabs(vel(r, ru)) + snd(P.gamma, r, ru, E));
        }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = jmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double lam = jmax(jmax(red[0], kEps), kEps);
    __syncthreads();
    double d = INFINITY;
    for (int e = threadIdx.x; e < Nel; e += kThreads)
        d = jmin(d, cfl * P.dx[e] / (double(2 * (P.np - 1) + 1) * lam));
    red[threadIdx.x] = d;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = jmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double dt = red[0];
    __syncthreads();
    return dt;
}

// ssp_rk3_step! 1D (SSP_RK3.jl:30-62): returns 0 ok, else the failing stage (4: final state)
__device__ int rk3_step(const Par1& P, double dt, double* t) {
#pragma clang fp contract(off)
    const int n = P.np * P.nel * kNeq1;
    const int tid = threadIdx.x;
    for (int k = tid; k < n; k += kThreads) P.u0[k] = P.u[k];
    __syncthreads();
    const double t0 = *t;
    residual(P, P.u);
    if (any_nonfinite(P.du, n)) return 1;
    for (int k = tid; k < n; k += kThreads) P.u[k] = P.u0[k] + dt * P.du[k];
    *t = t0 + dt;
    __syncthreads();
    residual(P, P.u);
    if (any_nonfinite(P.du, n)) return 2;
    for (int k = tid; k < n; k += kThreads) {
        P.u1[k] = P.u[k];
        P.u[k] = (3.0 / 4.0) * P.u0[k] + (1.0 / 4.0) * P.u1[k] + ((1.0 / 4.0) * dt) * P.du[k];
    }
    *t = t0 + dt / 2.0;
    __syncthreads();
    residual(P, P.u);
    if (any_nonfinite(P.du, n)) return 3;
    for (int k = tid; k < n; k += kThreads) {
        P.u1[k] = P.u[k];
        P.u[k] = (1.0 / 3.0) * P.u0[k] + (2.0 / 3.0) * P.u1[k] + ((2.0 / 3.0) * dt) * P.du[k];
    }
    *t = t0 + dt;
    __syncthreads();
    if (any_nonfinite(P.u, n)) return 4;
    return 0;
}

// One launch, one workgroup: mode 0 residual of u into du; mode 1 one step of fixed_dt; mode 2
// ssp_rk3! to t_final (SSP_RK3.jl:73-118).  Thread 0 keeps the step control.
__global__ void __launch_bounds__(kThreads) run_kernel(Par1 P) {
    __shared__ double red[kThreads];
    Ctl1& c = *P.ctl;
    const int mode = c.mode;
    if (mode == 0) {
        residual(P, P.u);
        return;
    }
    double t = c.t;
    if (mode == 1) {
        const int fs = rk3_step(P, c.fixed_dt, &t);
        if (threadIdx.x == 0) {
            c.t = t;
            c.fail_stage = fs;
            c.status = fs ? 1 : 0;
            c.n_steps = 1;
        }
        return;
    }
    int64_t n = 0;
    int status = 0, fs = 0;
    const double EPS10 = 10.0 * kEps;
    while (t < c.t_final - EPS10) {
        double step = c.fixed_dt > 0.0 ? c.fixed_dt : compute_dt(P, P.u, c.cfl, red);
        if (t + step > c.t_final) step = c.t_final - t;
        fs = rk3_step(P, step, &t);
        n += 1;
        if (fs) {
            status = 1;
            break;
        }
        if (n >= c.max_steps) {
            status = 2;
            break;
        }
    }
    if (threadIdx.x == 0) {
        c.t = t;
        c.n_steps = n;
        c.status = status;
        c.fail_stage = fs;
    }
}

__global__ void dt_kernel(Par1 P, double cfl, double* out) {
    __shared__ double red[kThreads];
    const double d = compute_dt(P, P.u, cfl, red);
    if (threadIdx.x == 0) *out = d;
}

#undef UU
#undef EE

}  // namespace d1
}  // namespace frf

// ---------------------------------------------------------------- C ABI

using frf::d1::Ctl1;
using frf::d1::Ops1;
using frf::d1::Par1;

struct frf1d_ctx {
    frf1d_problem pb;
    int np = 0, nel = 0;
    size_t n = 0;
    std::string err;
    hipStream_t stream = nullptr;
    double* dbuf = nullptr;     // u, u0, u1, du, g and the per-element / per-face scratch
    Ops1* dOps = nullptr;
    Ctl1* dCtl = nullptr;
    Ctl1* hCtl = nullptr;       // pinned
    double* dOut = nullptr;
    Par1 P{};
    Ops1 ops{};
    bool ops_set = false, geom_set = false, state_set = false;
    bool bc_set[2] = {false, false};   // Dirichlet ends with a state
};

static thread_local std::string g_err1;

static int32_t fail1(frf1d_ctx* c, int32_t code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf; else g_err1 = buf;
    return code;
}
#define HIPCHK1(c, expr)                                                                     \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) return fail1((c), FRF_E_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

extern "C" const char* frf1d_last_error(const frf1d_ctx* c) { return c ? c->err.c_str() : g_err1.c_str(); }

extern "C" int32_t frf1d_destroy(frf1d_ctx* c) {
    if (!c) return FRF_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->dbuf) (void)hipFree(c->dbuf);
    if (c->dOps) (void)hipFree(c->dOps);
    if (c->dCtl) (void)hipFree(c->dCtl);
    if (c->dOut) (void)hipFree(c->dOut);
    if (c->hCtl) (void)hipHostFree(c->hCtl);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FRF_OK;
}

extern "C" int32_t frf1d_create(const frf1d_problem* pb, frf1d_ctx** out) {
    if (!out) return FRF_E_ARG;
    *out = nullptr;
    if (!pb) return fail1(nullptr, FRF_E_ARG, "null problem");
    if (pb->p < 1 || pb->p > 5) return fail1(nullptr, FRF_E_UNSUPPORTED, "p must be in 1..5 (got %d)", pb->p);
    if (pb->nel < 1) return fail1(nullptr, FRF_E_ARG, "nel must be >= 1");
    if (pb->neq != 3) return fail1(nullptr, FRF_E_UNSUPPORTED, "the 1D device path implements Euler1D (neq = 3)");
    if ((pb->bc_left == FRF_BC_PERIODIC) != (pb->bc_right == FRF_BC_PERIODIC))
        return fail1(nullptr, FRF_E_ARG, "Periodic BC must be set on both ends");
    for (int32_t b : {pb->bc_left, pb->bc_right})
        if (b != FRF_BC_PERIODIC && b != FRF_BC_TRANSMISSIVE && b != FRF_BC_DIRICHLET)
            return fail1(nullptr, FRF_E_UNSUPPORTED, "unsupported BC %d", b);
    if (pb->flux != FRF_FLUX_RUSANOV) return fail1(nullptr, FRF_E_UNSUPPORTED, "the 1D device path implements the Rusanov flux");
    if (pb->precision != FRF_FP64) return fail1(nullptr, FRF_E_UNSUPPORTED, "the 1D device path runs in FP64");
    if (pb->capturing < FRF_CAPT_NULL || pb->capturing > FRF_CAPT_PERSSON_DD)
        return fail1(nullptr, FRF_E_ARG, "unknown capturing kind %d", pb->capturing);
    frf1d_ctx* c = new frf1d_ctx();
    c->pb = *pb;
    c->np = pb->p + 1;
    c->nel = pb->nel;
    c->n = (size_t)c->np * c->nel * 3;
    const size_t ne = (size_t)c->nel * 3;
    // u, u0, u1, du, g (n each); uL uR fL fR gL gR ghL ghR (ne each); sig eps dx J (nel each)
    const size_t total = 5 * c->n + 8 * ne + 4 * (size_t)c->nel;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->dbuf, total * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->dOps, sizeof(Ops1)) != hipSuccess || hipMalloc(&c->dCtl, sizeof(Ctl1)) != hipSuccess ||
        hipMalloc(&c->dOut, sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->hCtl, sizeof(Ctl1), hipHostMallocDefault) != hipSuccess) {
        g_err1 = "frf1d_create: device allocation failed";
        frf1d_destroy(c);
        return FRF_E_HIP;
    }
    double* b = c->dbuf;
    Par1& P = c->P;
    P.u = b; b += c->n;
    P.u0 = b; b += c->n;
    P.u1 = b; b += c->n;
    P.du = b; b += c->n;
    P.g = b; b += c->n;
    P.uL = b; b += ne; P.uR = b; b += ne; P.fL = b; b += ne; P.fR = b; b += ne;
    P.gL = b; b += ne; P.gR = b; b += ne; P.ghL = b; b += ne; P.ghR = b; b += ne;
    P.sig = b; b += c->nel;
    P.eps = b; b += c->nel;
    double* dx = b; b += c->nel;
    double* J = b;
    P.dx = dx;
    P.J = J;
    P.op = c->dOps;
    P.ctl = c->dCtl;
    P.np = c->np;
    P.nel = c->nel;
    P.periodic = pb->bc_left == FRF_BC_PERIODIC;
    P.dir_l = pb->bc_left == FRF_BC_DIRICHLET;
    P.dir_r = pb->bc_right == FRF_BC_DIRICHLET;
    P.capt = pb->capturing;
    P.gamma = pb->gamma;
    P.kappa = pb->kappa;
    P.s0 = pb->s0_factor * std::log10((double)(pb->p > 1 ? pb->p : 1));   // persson_s0
    P.eps_floor = pb->eps_floor;
    P.c_av = pb->c_av;
    *out = c;
    return FRF_OK;
}

extern "C" int32_t frf1d_set_boundary_state(frf1d_ctx* c, int32_t side, const double* U) {
    if (!c) return FRF_E_ARG;
    if (side < 0 || side > 1 || !U) return fail1(c, FRF_E_ARG, "frf1d_set_boundary_state: bad side or state");
    if (!(side == 0 ? c->P.dir_l : c->P.dir_r)) return fail1(c, FRF_E_ARG, "end %d is not a Dirichlet end", side);
    for (int q = 0; q < 3; ++q) {
        if (!std::isfinite(U[q])) return fail1(c, FRF_E_ARG, "non-finite Dirichlet state");
        (side == 0 ? c->P.ud_l : c->P.ud_r)[q] = U[q];
    }
    c->bc_set[side] = true;
    return FRF_OK;
}

extern "C" int32_t frf1d_set_operators(frf1d_ctx* c, const double* xi, const double* w, const double* D,
                                       const double* lL, const double* lR, const double* gLxi,
                                       const double* gRxi, const double* V) {
    if (!c) return FRF_E_ARG;
    if (!D || !lL || !lR || !gLxi || !gRxi || !V) return fail1(c, FRF_E_ARG, "null operator array");
    (void)xi;
    (void)w;
    const int n = c->np;
    Ops1 o{};
    for (int k = 0; k < n * n; ++k) o.D[k] = D[k];
    for (int j = 0; j < n; ++j) {
        o.lL[j] = lL[j];
        o.lR[j] = lR[j];
        o.gL[j] = gLxi[j];
        o.gR[j] = gRxi[j];
    }
    // LU with partial pivoting of V (column-major), the factorisation the sensor's V \ u uses
    for (int k = 0; k < n * n; ++k) o.LU[k] = V[k];
    for (int k = 0; k < n; ++k) {
        int p = k;
        double m = std::fabs(o.LU[k + n * k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(o.LU[i + n * k]) > m) { m = std::fabs(o.LU[i + n * k]); p = i; }
        if (!(m > 0.0)) return fail1(c, FRF_E_ARG, "Legendre Vandermonde is singular");
        o.piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(o.LU[k + n * j], o.LU[p + n * j]);
        for (int i = k + 1; i < n; ++i) {
            o.LU[i + n * k] /= o.LU[k + n * k];
            for (int j = k + 1; j < n; ++j) o.LU[i + n * j] -= o.LU[i + n * k] * o.LU[k + n * j];
        }
    }
    c->ops = o;
    HIPCHK1(c, hipMemcpy(c->dOps, &c->ops, sizeof(Ops1), hipMemcpyHostToDevice));
    c->ops_set = true;
    return FRF_OK;
}

extern "C" int32_t frf1d_set_geometry(frf1d_ctx* c, const double* xv) {
#pragma clang fp contract(off)
    if (!c) return FRF_E_ARG;
    if (!xv) return fail1(c, FRF_E_ARG, "null vertex array");
    std::vector<double> dx(c->nel), J(c->nel);
    for (int e = 0; e < c->nel; ++e) {
        if (!(xv[e + 1] > xv[e])) return fail1(c, FRF_E_ARG, "vertices must increase");
        dx[e] = xv[e + 1] - xv[e];   // Mesh1D.jl:44-47
        J[e] = dx[e] / 2.0;
    }
    HIPCHK1(c, hipMemcpy(const_cast<double*>(c->P.dx), dx.data(), c->nel * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK1(c, hipMemcpy(const_cast<double*>(c->P.J), J.data(), c->nel * sizeof(double), hipMemcpyHostToDevice));
    c->geom_set = true;
    return FRF_OK;
}

extern "C" int32_t frf1d_upload_state(frf1d_ctx* c, const double* u, int64_t len, double t) {
    if (!c) return FRF_E_ARG;
    if (!u || len != (int64_t)c->n) return fail1(c, FRF_E_ARG, "state length %lld != %zu", (long long)len, c->n);
    HIPCHK1(c, hipMemcpy(c->P.u, u, c->n * sizeof(double), hipMemcpyHostToDevice));
    c->hCtl->t = t;
    c->state_set = true;
    return FRF_OK;
}

extern "C" int32_t frf1d_download_state(frf1d_ctx* c, double* u, int64_t len, double* t) {
    if (!c) return FRF_E_ARG;
    if (!u || len != (int64_t)c->n) return fail1(c, FRF_E_ARG, "state length %lld != %zu", (long long)len, c->n);
    HIPCHK1(c, hipStreamSynchronize(c->stream));
    HIPCHK1(c, hipMemcpy(u, c->P.u, c->n * sizeof(double), hipMemcpyDeviceToHost));
    if (t) *t = c->hCtl->t;
    return FRF_OK;
}

static int32_t ready1(frf1d_ctx* c) {
    if (!c) return FRF_E_ARG;
    if (!c->ops_set) return fail1(c, FRF_E_STATE, "operators not set");
    if (!c->geom_set) return fail1(c, FRF_E_STATE, "geometry not set");
    if (!c->state_set) return fail1(c, FRF_E_STATE, "no state uploaded");
    if ((c->P.dir_l && !c->bc_set[0]) || (c->P.dir_r && !c->bc_set[1]))
        return fail1(c, FRF_E_STATE, "a Dirichlet end has no state (frf1d_set_boundary_state)");
    return FRF_OK;
}

static int32_t launch1(frf1d_ctx* c, int mode, double t_final, double cfl, double fixed_dt, int64_t max_steps) {
    Ctl1& h = *c->hCtl;
    h.t_final = t_final;
    h.cfl = cfl;
    h.fixed_dt = fixed_dt;
    h.max_steps = max_steps;
    h.n_steps = 0;
    h.status = 0;
    h.fail_stage = 0;
    h.mode = mode;
    HIPCHK1(c, hipMemcpyAsync(c->dCtl, c->hCtl, sizeof(Ctl1), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(frf::d1::run_kernel, dim3(1), dim3(frf::d1::kThreads), 0, c->stream, c->P);
    HIPCHK1(c, hipGetLastError());
    HIPCHK1(c, hipMemcpyAsync(c->hCtl, c->dCtl, sizeof(Ctl1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK1(c, hipStreamSynchronize(c->stream));
    return FRF_OK;
}

extern "C" int32_t frf1d_residual(frf1d_ctx* c, double* du, int64_t len) {
    int32_t r = ready1(c);
    if (r) return r;
    if (!du || len != (int64_t)c->n) return fail1(c, FRF_E_ARG, "du length %lld != %zu", (long long)len, c->n);
    if ((r = launch1(c, 0, 0.0, 0.0, 0.0, 1))) return r;
    HIPCHK1(c, hipMemcpy(du, c->P.du, c->n * sizeof(double), hipMemcpyDeviceToHost));
    return FRF_OK;
}

extern "C" int32_t frf1d_rk3_step(frf1d_ctx* c, double dt, int32_t* status) {
    int32_t r = ready1(c);
    if (r) return r;
    if (!(dt > 0.0)) return fail1(c, FRF_E_ARG, "dt must be > 0");
    if ((r = launch1(c, 1, INFINITY, 0.0, dt, 1))) return r;
    if (status) *status = c->hCtl->status ? FRF_STATUS_DIVERGED : FRF_STATUS_OK;
    return FRF_OK;
}

extern "C" int32_t frf1d_run(frf1d_ctx* c, double t_final, double cfl, double fixed_dt, int64_t max_steps,
                             frf_run_result* res) {
    int32_t r = ready1(c);
    if (r) return r;
    if (!res) return fail1(c, FRF_E_ARG, "null result");
    if (max_steps < 1) return fail1(c, FRF_E_ARG, "max_steps must be >= 1");
    if (!(fixed_dt > 0.0) && !(cfl > 0.0)) return fail1(c, FRF_E_ARG, "cfl must be > 0 for CFL steps");
    if ((r = launch1(c, 2, t_final, cfl, fixed_dt > 0.0 ? fixed_dt : 0.0, max_steps))) return r;
    const Ctl1& h = *c->hCtl;
    res->status = h.status;
    res->fail_stage = h.fail_stage;
    res->n_steps = h.n_steps;
    res->t = h.t;
    return FRF_OK;
}

extern "C" int32_t frf1d_compute_dt(frf1d_ctx* c, double cfl, double* dt) {
    int32_t r = ready1(c);
    if (r) return r;
    if (!dt) return fail1(c, FRF_E_ARG, "null dt");
    hipLaunchKernelGGL(frf::d1::dt_kernel, dim3(1), dim3(frf::d1::kThreads), 0, c->stream, c->P, cfl, c->dOut);
    HIPCHK1(c, hipGetLastError());
    HIPCHK1(c, hipMemcpyAsync(dt, c->dOut, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK1(c, hipStreamSynchronize(c->stream));
    return FRF_OK;
}
