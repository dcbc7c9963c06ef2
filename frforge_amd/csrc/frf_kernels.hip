// gfx950 kernels for the 2D FR residual and SSP-RK3 stage update.
//
// Mapping.  Each element of degree p has Np = p+1 solution-point rows and columns; the
// element is owned by Np consecutive threads ("line threads").  Thread t of an element first
// acts as column i = t (eta sweep) and then as row j = t (xi sweep).  Everything along a line
// (physical flux, contravariant flux, derivative D*F, Lagrange traces, g_DG corrections) is
// computed in registers; the only exchange inside an element is the eta contribution, passed
// through LDS from the column owners to the row owners.
//
// Faces.  A line thread computes the numerical flux on both ends of its line, loading the
// neighbour's matching line from global memory (mostly L2 hits: the neighbour was just read by
// the adjacent lanes or by a block scheduled on the same XCD a moment earlier).  Both elements
// that share a face evaluate its flux from identical inputs in identical order (left trace,
// right trace, the left element's normal), so the two copies agree bit for bit.  No atomics,
// no face buffers in HBM.
//
// Stage fusion.  The residual, the SSP-RK3 combine (in FP64), and the non-finite checks of the
// reference's ssp_rk3_step! run in one kernel per stage; stage 3 updates u^n in place.
//
// Metrics.  Affine meshes carry their metric terms as per-column and per-row tables that
// reproduce the reference's D-derived values bit for bit (MeshMetrics2D.jl:82-218); the Jacobian
// and the face normals are rebuilt from them per point, so no per-point metric array is read.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "frf_kernels.h"
#include "frf_launch.h"
#include "frf_physics.h"

namespace frf {

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs.  Remap so that each
// XCD walks a contiguous range of element strips: a strip's y-neighbours (nx elements away)
// are then processed on the same XCD shortly before or after it and are served by its L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    const int per = nblk >> 3, rem = nblk & 7;
    const int x = bid & 7, k = bid >> 3;
    return x * per + (x < rem ? x : rem) + k;
}

template <int NP, typename R>
__device__ __forceinline__ void load_row(const double* __restrict__ u, int nel, int e, int j,
                                         R out[NP][kNeq]) {
#pragma unroll
    for (int c = 0; c < kNeq; ++c) {
        const double* p = u + ((size_t)c * nel + e) * (NP * NP) + j * NP;
#pragma unroll
        for (int i = 0; i < NP; ++i) out[i][c] = R(p[i]);
    }
}

template <int NP, typename R>
__device__ __forceinline__ void load_col(const double* __restrict__ u, int nel, int e, int i,
                                         R out[NP][kNeq]) {
#pragma unroll
    for (int c = 0; c < kNeq; ++c) {
        const double* p = u + ((size_t)c * nel + e) * (NP * NP) + i;
#pragma unroll
        for (int j = 0; j < NP; ++j) out[j][c] = R(p[j * NP]);
    }
}

// Lagrange trace of a line: sum_k w[k] * v[k][c], accumulated from zero in index order.
template <int NP, typename R>
__device__ __forceinline__ void trace(const double* w, const R v[NP][kNeq], R out[kNeq]) {
#pragma unroll
    for (int c = 0; c < kNeq; ++c) {
        R s = R(0);
#pragma unroll
        for (int k = 0; k < NP; ++k) s += R(w[k]) * v[k][c];
        out[c] = s;
    }
}

// Unit normal and surface Jacobian from the two tangential metric components of a face.
// East (+xi):  n ~ ( y_eta, -x_eta);  west: the negative;  (MeshMetrics2D.jl:140-175)
// North (+eta): n ~ (-y_xi,  x_xi);   south: the negative.
template <typename R>
__device__ __forceinline__ R face_s(R a, R b) {
    R s = rsqrt_(a * a + b * b);
    return rmax(s, R(kEps64));
}

// SSP-RK3 combine in FP64 (SSP_RK3.jl:149,154,159), reference operation order, no FMA.
__device__ __forceinline__ double rk_combine(int stage, double u0, double uk, double du, double dt) {
#pragma clang fp contract(off)
    if (stage == 1) return u0 + dt * du;
    if (stage == 2) return (3.0 / 4.0) * u0 + (1.0 / 4.0) * uk + ((1.0 / 4.0) * dt) * du;
    return (1.0 / 3.0) * u0 + (2.0 / 3.0) * uk + ((2.0 / 3.0) * dt) * du;
}

// BR0 interior penalty and averaged flux (SharedAV.jl:29-42)
template <typename R>
__device__ __forceinline__ R br0_tau(double em, double ep, double hm, double hp, int p) {
    double ebar = 0.5 * (em + ep);
    double hbar = 0.5 * (hm + hp);
    return R(2.0 * ebar * double(p + 1) / (hbar > kEps64 ? hbar : kEps64));
}
template <typename R>
__device__ __forceinline__ R br0_avg(R gm, R gp, R um, R up, R tau) {
    return R(0.5) * (gm + gp) - tau * (up - um);
}

template <int NP, typename R, int FLUX, int CAPT, int EPB>
__global__ void __launch_bounds__(EPB* NP) stage_kernel(StageParams P, OpsT<NP> op) {
    using L = TabLayout<NP>;
    constexpr int CT = L::STRIDE;
    constexpr int NB = (CAPT != 0) ? 2 : 1;
    __shared__ R sB[NB][EPB * kNeq * NP * NP];

    // Skip after a failure in an earlier stage of this step.  The test reads only values written
    // by earlier launches (this stage reports P.stage or 4, never less), so it is uniform.
    if (P.stage > 0) {
        if (P.ctl->done != 0 || P.ctl->fail_stage < P.stage) return;
    }
    const int tid = threadIdx.x;
    const int le = tid / NP;
    const int t = tid - le * NP;
    const int blk = xcd_remap(blockIdx.x, P.nblk);
    const int e = blk * EPB + le;
    const bool valid = e < P.nel;
    const int nx = P.nx, ny = P.ny, nel = P.nel;
    const R g = R(P.gamma);
    const int jx = valid ? e % nx : 0;
    const int jy = valid ? e / nx : 0;
    const double* ct = P.ctab + (size_t)jx * CT;
    const double* rt = P.rtab + (size_t)jy * CT;
    double eps_e = 0.0;
    if constexpr (CAPT != 0) {
        if (valid) eps_e = P.eps[e];
    }
    R* myB = &sB[0][((size_t)le * kNeq) * NP * NP];

    // ------------------------------------------------------------ column pass (i = t)
    if (valid) {
        const int i = t;
        R uc[NP][kNeq];
        load_col<NP>(P.uin, nel, e, i, uc);
        R Gt[NP][kNeq];
        const R xxi = R(ct[L::XI + i]);
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            Prim<R> q = primitives(uc[j], g);
            R F[4], G[4];
            phys_flux(uc[j], q, F, G);
            const R yxi = R(rt[L::CROSS + j * NP + i]);
#pragma unroll
            for (int c = 0; c < kNeq; ++c) Gt[j][c] = -yxi * F[c] + xxi * G[c];
        }
        R trS[kNeq], trN[kNeq], gS[kNeq], gN[kNeq];
        trace<NP>(op.lL, uc, trS);
        trace<NP>(op.lR, uc, trN);
        trace<NP>(op.lL, Gt, gS);
        trace<NP>(op.lR, Gt, gN);

        // neighbours along eta
        const int eS = (jy > 0) ? e - nx : (P.bcy == 0 ? e + (ny - 1) * nx : -1);
        const int eN = (jy < ny - 1) ? e + nx : (P.bcy == 0 ? e - (ny - 1) * nx : -1);
        R fhS[kNeq], fhN[kNeq];
        // BR0 pieces (only used when CAPT == 1)
        R gy[NP][kNeq];
        R gyS[kNeq], gyN[kNeq], ghS[kNeq], ghN[kNeq];
        const double Jy = rt[L::JAC], dy = rt[L::DELTA];
        if constexpr (CAPT == 1) {
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int c = 0; c < kNeq; ++c) {
                    R s = R(0);
#pragma unroll
                    for (int k = 0; k < NP; ++k) s += R(op.D[j * NP + k]) * uc[k][c];
                    gy[j][c] = R(eps_e) * s / R(Jy);
                }
            trace<NP>(op.lL, gy, gyS);
            trace<NP>(op.lR, gy, gyN);
        }
        {   // south face: flux with the lower element's north normal
            if (eS >= 0) {
                R un[NP][kNeq];
                load_col<NP>(P.uin, nel, eS, i, un);
                R trNn[kNeq];
                trace<NP>(op.lR, un, trNn);
                const int jyS = eS / nx;
                const double* rtS = P.rtab + (size_t)jyS * CT;
                R a = R(rtS[L::FN + i]), b = R(ct[L::FN + i]);   // y_xi, x_xi at N face
                R s = face_s(a, b);
                R fh[4];
                iflux<FLUX>(trNn, trS, -a / s, b / s, g, fh);
#pragma unroll
                for (int c = 0; c < kNeq; ++c) fhS[c] = fh[c] * s;
                if constexpr (CAPT == 1) {
                    const double epsn = P.eps[eS];
                    const double Jyn = rtS[L::JAC], dyn = rtS[L::DELTA];
                    R gyn[NP][kNeq];
#pragma unroll
                    for (int j = 0; j < NP; ++j)
#pragma unroll
                        for (int c = 0; c < kNeq; ++c) {
                            R sum = R(0);
#pragma unroll
                            for (int k = 0; k < NP; ++k) sum += R(op.D[j * NP + k]) * un[k][c];
                            gyn[j][c] = R(epsn) * sum / R(Jyn);
                        }
                    R gyNn[kNeq];
                    trace<NP>(op.lR, gyn, gyNn);
                    const R tau = br0_tau<R>(epsn, eps_e, dyn, dy, P.p);
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghS[c] = br0_avg(gyNn[c], gyS[c], trNn[c], trS[c], tau);
                }
            } else {
                R a = R(rt[L::FS + i]), b = R(ct[L::FS + i]);    // y_xi, x_xi at S face
                R s = face_s(a, b);
                R nxo = a / s, nyo = -b / s;                     // outward south normal
                R fh[4];
                iflux<FLUX>(trS, trS, -nxo, -nyo, g, fh);
#pragma unroll
                for (int c = 0; c < kNeq; ++c) fhS[c] = fh[c] * s;
                if constexpr (CAPT == 1) {
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghS[c] = gyS[c];
                }
            }
        }
        {   // north face: own north normal
            R a = R(rt[L::FN + i]), b = R(ct[L::FN + i]);
            R s = face_s(a, b);
            R nxo = -a / s, nyo = b / s;
            R fh[4];
            if (eN >= 0) {
                R un[NP][kNeq];
                load_col<NP>(P.uin, nel, eN, i, un);
                R trSn[kNeq];
                trace<NP>(op.lL, un, trSn);
                iflux<FLUX>(trN, trSn, nxo, nyo, g, fh);
                if constexpr (CAPT == 1) {
                    const double epsn = P.eps[eN];
                    const double* rtN = P.rtab + (size_t)(eN / nx) * CT;
                    const double Jyn = rtN[L::JAC], dyn = rtN[L::DELTA];
                    R gyn[NP][kNeq];
#pragma unroll
                    for (int j = 0; j < NP; ++j)
#pragma unroll
                        for (int c = 0; c < kNeq; ++c) {
                            R sum = R(0);
#pragma unroll
                            for (int k = 0; k < NP; ++k) sum += R(op.D[j * NP + k]) * un[k][c];
                            gyn[j][c] = R(epsn) * sum / R(Jyn);
                        }
                    R gySn[kNeq];
                    trace<NP>(op.lL, gyn, gySn);
                    const R tau = br0_tau<R>(eps_e, epsn, dy, dyn, P.p);
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghN[c] = br0_avg(gyN[c], gySn[c], trN[c], trSn[c], tau);
                }
            } else {
                iflux<FLUX>(trN, trN, nxo, nyo, g, fh);
                if constexpr (CAPT == 1) {
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghN[c] = gyN[c];
                }
            }
#pragma unroll
            for (int c = 0; c < kNeq; ++c) fhN[c] = fh[c] * s;
        }
        // eta contribution / J, written for the row owners
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const R J = xxi * R(rt[L::XI + j]) - R(ct[L::CROSS + j * NP + i]) * R(rt[L::CROSS + j * NP + i]);
#pragma unroll
            for (int c = 0; c < kNeq; ++c) {
                R vol = R(0);
#pragma unroll
                for (int k = 0; k < NP; ++k) vol += R(op.D[j * NP + k]) * Gt[k][c];
                R corr = (fhS[c] - gS[c]) * R(op.gL[j]) + (fhN[c] - gN[c]) * R(op.gR[j]);
                myB[(c * NP + j) * NP + i] = (vol + corr) / J;
            }
        }
        if constexpr (CAPT == 1) {
            R* myBy = &sB[1][((size_t)le * kNeq) * NP * NP];
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int c = 0; c < kNeq; ++c) {
                    R vol = R(0);
#pragma unroll
                    for (int k = 0; k < NP; ++k) vol += R(op.D[j * NP + k]) * gy[k][c];
                    R corr = (ghS[c] - gyS[c]) * R(op.gL[j]) + (ghN[c] - gyN[c]) * R(op.gR[j]);
                    myBy[(c * NP + j) * NP + i] = (vol + corr) / R(Jy);
                }
        }
        if constexpr (CAPT == 2) {
            R* myBy = &sB[1][((size_t)le * kNeq) * NP * NP];
            const R sy = R(eps_e / (Jy * Jy));
#pragma unroll
            for (int c = 0; c < kNeq; ++c) {
                R v[NP];
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    R s = R(0);
#pragma unroll
                    for (int k = 0; k < NP; ++k) s += R(op.D[j * NP + k]) * uc[k][c];
                    v[j] = s;
                }
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    R s = R(0);
#pragma unroll
                    for (int k = 0; k < NP; ++k) s += R(op.D[j * NP + k]) * v[k];
                    myBy[(c * NP + j) * NP + i] = sy * s;
                }
            }
        }
    }
    __syncthreads();

    // ------------------------------------------------------------ row pass (j = t)
    bool bad = false;
    bool bad_u = false;
    if (valid) {
        const int j = t;
        R ur[NP][kNeq];
        load_row<NP>(P.uin, nel, e, j, ur);
        R Ft[NP][kNeq];
        const R yeta = R(rt[L::XI + j]);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            Prim<R> q = primitives(ur[i], g);
            R F[4], G[4];
            phys_flux(ur[i], q, F, G);
            const R xeta = R(ct[L::CROSS + j * NP + i]);
#pragma unroll
            for (int c = 0; c < kNeq; ++c) Ft[i][c] = yeta * F[c] - xeta * G[c];
        }
        R trW[kNeq], trE[kNeq], fL[kNeq], fR[kNeq];
        trace<NP>(op.lL, ur, trW);
        trace<NP>(op.lR, ur, trE);
        trace<NP>(op.lL, Ft, fL);
        trace<NP>(op.lR, Ft, fR);

        const int eW = (jx > 0) ? e - 1 : (P.bcx == 0 ? e + nx - 1 : -1);
        const int eE = (jx < nx - 1) ? e + 1 : (P.bcx == 0 ? e - (nx - 1) : -1);
        R fhW[kNeq], fhE[kNeq];
        R gx[NP][kNeq];
        R gxW[kNeq], gxE[kNeq], ghW[kNeq], ghE[kNeq];
        const double Jx = ct[L::JAC], dx = ct[L::DELTA];
        if constexpr (CAPT == 1) {
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int c = 0; c < kNeq; ++c) {
                    R s = R(0);
#pragma unroll
                    for (int k = 0; k < NP; ++k) s += R(op.D[i * NP + k]) * ur[k][c];
                    gx[i][c] = R(eps_e) * s / R(Jx);
                }
            trace<NP>(op.lL, gx, gxW);
            trace<NP>(op.lR, gx, gxE);
        }
        {   // west face: flux with the left element's east normal
            if (eW >= 0) {
                R un[NP][kNeq];
                load_row<NP>(P.uin, nel, eW, j, un);
                R trEn[kNeq];
                trace<NP>(op.lR, un, trEn);
                const double* ctW = P.ctab + (size_t)(eW % nx) * CT;
                R a = R(rt[L::FE + j]), b = R(ctW[L::FE + j]);   // y_eta, x_eta at E face
                R s = face_s(a, b);
                R fh[4];
                iflux<FLUX>(trEn, trW, a / s, -b / s, g, fh);
#pragma unroll
                for (int c = 0; c < kNeq; ++c) fhW[c] = fh[c] * s;
                if constexpr (CAPT == 1) {
                    const double epsn = P.eps[eW];
                    const double Jxn = ctW[L::JAC], dxn = ctW[L::DELTA];
                    R gxn[NP][kNeq];
#pragma unroll
                    for (int i = 0; i < NP; ++i)
#pragma unroll
                        for (int c = 0; c < kNeq; ++c) {
                            R sum = R(0);
#pragma unroll
                            for (int k = 0; k < NP; ++k) sum += R(op.D[i * NP + k]) * un[k][c];
                            gxn[i][c] = R(epsn) * sum / R(Jxn);
                        }
                    R gxEn[kNeq];
                    trace<NP>(op.lR, gxn, gxEn);
                    const R tau = br0_tau<R>(epsn, eps_e, dxn, dx, P.p);
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghW[c] = br0_avg(gxEn[c], gxW[c], trEn[c], trW[c], tau);
                }
            } else {
                R a = R(rt[L::FW + j]), b = R(ct[L::FW + j]);    // y_eta, x_eta at W face
                R s = face_s(a, b);
                R nxo = -a / s, nyo = b / s;                     // outward west normal
                R fh[4];
                iflux<FLUX>(trW, trW, -nxo, -nyo, g, fh);
#pragma unroll
                for (int c = 0; c < kNeq; ++c) fhW[c] = fh[c] * s;
                if constexpr (CAPT == 1) {
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghW[c] = gxW[c];
                }
            }
        }
        {   // east face: own east normal
            R a = R(rt[L::FE + j]), b = R(ct[L::FE + j]);
            R s = face_s(a, b);
            R nxo = a / s, nyo = -b / s;
            R fh[4];
            if (eE >= 0) {
                R un[NP][kNeq];
                load_row<NP>(P.uin, nel, eE, j, un);
                R trWn[kNeq];
                trace<NP>(op.lL, un, trWn);
                iflux<FLUX>(trE, trWn, nxo, nyo, g, fh);
                if constexpr (CAPT == 1) {
                    const double epsn = P.eps[eE];
                    const double* ctE = P.ctab + (size_t)(eE % nx) * CT;
                    const double Jxn = ctE[L::JAC], dxn = ctE[L::DELTA];
                    R gxn[NP][kNeq];
#pragma unroll
                    for (int i = 0; i < NP; ++i)
#pragma unroll
                        for (int c = 0; c < kNeq; ++c) {
                            R sum = R(0);
#pragma unroll
                            for (int k = 0; k < NP; ++k) sum += R(op.D[i * NP + k]) * un[k][c];
                            gxn[i][c] = R(epsn) * sum / R(Jxn);
                        }
                    R gxWn[kNeq];
                    trace<NP>(op.lL, gxn, gxWn);
                    const R tau = br0_tau<R>(eps_e, epsn, dx, dxn, P.p);
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghE[c] = br0_avg(gxE[c], gxWn[c], trE[c], trWn[c], tau);
                }
            } else {
                iflux<FLUX>(trE, trE, nxo, nyo, g, fh);
                if constexpr (CAPT == 1) {
#pragma unroll
                    for (int c = 0; c < kNeq; ++c) ghE[c] = gxE[c];
                }
            }
#pragma unroll
            for (int c = 0; c < kNeq; ++c) fhE[c] = fh[c] * s;
        }

        // du = ((0 - xi/J) - eta/J) [+ AV_x] [+ AV_y]   (Residual2D.jl:371,393; PerssonAV2D.jl:373,392)
        R du[NP][kNeq];
        const R* myBc = myB;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const R J = R(ct[L::XI + i]) * yeta - R(ct[L::CROSS + j * NP + i]) * R(rt[L::CROSS + j * NP + i]);
#pragma unroll
            for (int c = 0; c < kNeq; ++c) {
                R vol = R(0);
#pragma unroll
                for (int k = 0; k < NP; ++k) vol += R(op.D[i * NP + k]) * Ft[k][c];
                R corr = (fhW[c] - fL[c]) * R(op.gL[i]) + (fhE[c] - fR[c]) * R(op.gR[i]);
                R a = (vol + corr) / J;
                du[i][c] = -a - myBc[(c * NP + j) * NP + i];
            }
        }
        if constexpr (CAPT == 1) {
            const R* myBy = &sB[1][((size_t)le * kNeq) * NP * NP];
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int c = 0; c < kNeq; ++c) {
                    R vol = R(0);
#pragma unroll
                    for (int k = 0; k < NP; ++k) vol += R(op.D[i * NP + k]) * gx[k][c];
                    R corr = (ghW[c] - gxW[c]) * R(op.gL[i]) + (ghE[c] - gxE[c]) * R(op.gR[i]);
                    du[i][c] = (du[i][c] + (vol + corr) / R(Jx)) + myBy[(c * NP + j) * NP + i];
                }
        }
        if constexpr (CAPT == 2) {
            if (eps_e > 0.0) {
                const R* myBy = &sB[1][((size_t)le * kNeq) * NP * NP];
                const R sx = R(eps_e / (Jx * Jx));
#pragma unroll
                for (int c = 0; c < kNeq; ++c) {
                    R v[NP];
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        R s = R(0);
#pragma unroll
                        for (int k = 0; k < NP; ++k) s += R(op.D[i * NP + k]) * ur[k][c];
                        v[i] = s;
                    }
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        R s = R(0);
#pragma unroll
                        for (int k = 0; k < NP; ++k) s += R(op.D[i * NP + k]) * v[k];
                        du[i][c] = (du[i][c] + sx * s) + myBy[(c * NP + j) * NP + i];
                    }
                }
            }
        }

        // stage combine / output
        const int stage = P.stage;
        const double dt = stage > 0 ? P.ctl->dt : 0.0;
#pragma unroll
        for (int c = 0; c < kNeq; ++c) {
            const size_t base = ((size_t)c * nel + e) * (NP * NP) + j * NP;
            double outv[NP];
            if (stage == 0) {
#pragma unroll
                for (int i = 0; i < NP; ++i) outv[i] = double(du[i][c]);
            } else {
                const double* pin = P.uin + base;
                const double* p0 = P.u0 + base;
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const double d = double(du[i][c]);
                    bad |= !isfinite(d);
                    outv[i] = rk_combine(stage, p0[i], pin[i], d, dt);
                    bad_u |= !isfinite(outv[i]);
                }
            }
            double* po = P.uout + base;
#pragma unroll
            for (int i = 0; i < NP; ++i) po[i] = outv[i];
        }
    }
    if (P.stage > 0) {
        const int any_bad = __syncthreads_or(bad ? 1 : 0);
        const int any_bad_u = (P.stage == 3) ? __syncthreads_or(bad_u ? 1 : 0) : 0;
        if (tid == 0 && (any_bad || any_bad_u)) atomicMin(&P.ctl->fail_stage, any_bad ? P.stage : 4);
    }
}

// Per-element wave speed, Persson sensor and element viscosity (PerssonAV2D.jl:20-109),
// plus the running maximum wave speed for the CFL step (Euler2D.jl:92-104).
template <int NP, typename R, int EPB>
__global__ void __launch_bounds__(EPB* NP) elem_kernel(ElemParams P, OpsT<NP> op, int accum_lambda,
                                                       int check_ctl) {
    using L = TabLayout<NP>;
    constexpr int CT = L::STRIDE;
    __shared__ R sT[EPB][NP][NP];
    __shared__ R sE[EPB][NP][2];
    __shared__ double sL[EPB * NP];
    if (check_ctl && (P.ctl->done != 0 || P.ctl->fail_stage < kNoFail)) return;
    const int tid = threadIdx.x;
    const int le = tid / NP;
    const int t = tid - le * NP;
    const int blk = xcd_remap(blockIdx.x, P.nblk);
    const int e = blk * EPB + le;
    const bool valid = e < P.nel;
    double lam = 0.0;
    if (valid) {
        const int j = t;
        double ur[NP][kNeq];
        load_row<NP>(P.u, P.nel, e, j, ur);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            double s = point_wavespeed<double>(ur[i], P.gamma);
            lam = (s > lam || s != s) ? s : lam;   // NaN-propagating max
        }
        if (P.want_sigma) {
#pragma unroll
            for (int a = 0; a < NP; ++a) {
                R s = R(0);
#pragma unroll
                for (int i = 0; i < NP; ++i) s += R(op.Vinv[a * NP + i]) * R(ur[i][0]);
                sT[le][a][j] = s;
            }
        }
    }
    sL[tid] = lam;
    __syncthreads();
    if (valid && P.want_sigma) {
        const int b = t;
        R et = R(0), eh = R(0), ecol = R(0);
#pragma unroll
        for (int a = 0; a < NP; ++a) {
            R s = R(0);
#pragma unroll
            for (int jj = 0; jj < NP; ++jj) s += sT[le][a][jj] * R(op.Vinv[b * NP + jj]);
            et += s * s;
            if (a == NP - 1) eh += s * s;
            else ecol += s * s;
        }
        if (b == NP - 1) eh += ecol;
        sE[le][b][0] = et;
        sE[le][b][1] = eh;
    }
    __syncthreads();
    if (valid && t == 0) {
        double lm = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            double s = sL[le * NP + k];
            lm = (s > lm || s != s) ? s : lm;
        }
        lm = (lm > kEps64 || lm != lm) ? lm : kEps64;
        if (P.want_sigma) {
            R etot = R(0), ehigh = R(0);
#pragma unroll
            for (int b = 0; b < NP; ++b) {
                etot += sE[le][b][0];
                ehigh += sE[le][b][1];
            }
            etot += R(P.eps_floor);
            R se;
            if constexpr (sizeof(R) == 8) {
                se = log10(ehigh / etot + P.eps_floor);
                se = 1.0 / (1.0 + exp(-P.kappa * (se - P.s0)));
            } else {
                se = log10f(ehigh / etot + float(P.eps_floor));
                se = 1.0f / (1.0f + expf(-float(P.kappa) * (se - float(P.s0))));
            }
            const int jx = e % P.nx, jy = e / P.nx;
            const double dx = P.ctab[(size_t)jx * CT + L::DELTA];
            const double dy = P.rtab[(size_t)jy * CT + L::DELTA];
            const double h = dx < dy ? dx : dy;
            const double sigma = double(se);
            const int pp = P.p > 1 ? P.p : 1;
            double ev = 0.0;
            if (sigma > 0.0) ev = P.c_av * sigma * (h / double(pp)) * lm;
            if (P.sigma) P.sigma[e] = sigma;
            if (P.eps) P.eps[e] = ev;
        }
        sL[le * NP] = lm;
    }
    if (accum_lambda) {
        __syncthreads();
        if (tid == 0) {
            double m = 0.0;
            for (int k = 0; k < EPB; ++k) {
                if (blk * EPB + k >= P.nel) break;
                double s = sL[k * NP];
                m = (s > m || s != s) ? s : m;
            }
            atomicMax(&P.ctl->lam_bits, (unsigned long long)__double_as_longlong(m));
        }
    }
}

// Step start: loop test, CFL or fixed step, last-step truncation (SSP_RK3.jl:185-189, 126-133).
__global__ void dt_kernel(DevCtl* c) {
    if (c->done != 0) return;
    const double t = c->t;
    if (!(t < c->t_final - 10.0 * kEps64)) {
        c->done = 1;
        return;
    }
    double step;
    if (c->fixed_dt > 0.0) {
        step = c->fixed_dt;
    } else {
        double lam = __longlong_as_double((long long)c->lam_bits);
        lam = (lam > kEps64 || lam != lam) ? lam : kEps64;
        step = c->cfl * c->hmin / (double(2 * c->p + 1) * lam);
    }
    c->lam_bits = 0ull;
    if (t + step > c->t_final) step = c->t_final - t;
    c->dt = step;
    c->t0 = t;
}

// Step end: step count, state time as the reference leaves it, status (SSP_RK3.jl:190-208).
__global__ void finish_kernel(DevCtl* c) {
    if (c->done != 0) return;
    c->n_steps += 1;
    if (c->fail_stage < kNoFail) {
        const int fs = c->fail_stage;
        c->status = 1;
        c->t = (fs == 1) ? c->t0 : (fs == 3 ? c->t0 + c->dt / 2.0 : c->t0 + c->dt);
        c->done = 1;
        return;
    }
    c->t = c->t0 + c->dt;
    if (c->n_steps >= c->max_steps) {
        c->status = 2;
        c->done = 1;
    }
}

// ---------------------------------------------------------------- host-side launchers

template <int NP>
static OpsT<NP> make_ops(const HostOps& h) {
    OpsT<NP> o;
    for (int a = 0; a < NP * NP; ++a) {
        o.D[a] = h.D[a];
        o.Vinv[a] = h.Vinv[a];
    }
    for (int a = 0; a < NP; ++a) {
        o.lL[a] = h.lL[a];
        o.lR[a] = h.lR[a];
        o.gL[a] = h.gLxi[a];
        o.gR[a] = h.gRxi[a];
    }
    return o;
}

template <int NP, typename R, int CAPT>
constexpr int stage_epb() {
    // elements per block: keep the LDS exchange at <= 32 KiB and blocks at 128-384 threads
    return (NP <= 2) ? 64 : ((NP <= 4) ? (sizeof(R) == 8 && CAPT != 0 ? 32 : 64) : 32);
}

template <int NP, typename R, int FLUX, int CAPT>
static hipError_t launch_stage_t(const StageParams& P0, const HostOps& h, hipStream_t s) {
    constexpr int EPB = stage_epb<NP, R, CAPT>();
    StageParams P = P0;
    P.nblk = (P.nel + EPB - 1) / EPB;
    hipLaunchKernelGGL((stage_kernel<NP, R, FLUX, CAPT, EPB>), dim3(P.nblk), dim3(EPB * NP), 0, s,
                       P, make_ops<NP>(h));
    return hipGetLastError();
}

template <int NP, typename R>
static hipError_t launch_stage_fc(int flux, int capt, const StageParams& P, const HostOps& h,
                                  hipStream_t s) {
    if (flux == 0) {
        if (capt == 0) return launch_stage_t<NP, R, 0, 0>(P, h, s);
        if (capt == 1) return launch_stage_t<NP, R, 0, 1>(P, h, s);
        return launch_stage_t<NP, R, 0, 2>(P, h, s);
    }
    if (capt == 0) return launch_stage_t<NP, R, 1, 0>(P, h, s);
    if (capt == 1) return launch_stage_t<NP, R, 1, 1>(P, h, s);
    return launch_stage_t<NP, R, 1, 2>(P, h, s);
}

template <int NP>
static hipError_t launch_stage_np(int prec, int flux, int capt, const StageParams& P,
                                  const HostOps& h, hipStream_t s) {
    if (prec == 0) return launch_stage_fc<NP, double>(flux, capt, P, h, s);
    return launch_stage_fc<NP, float>(flux, capt, P, h, s);
}

hipError_t launch_stage(int np, int prec, int flux, int capt, const StageParams& P,
                        const HostOps& h, hipStream_t s) {
    switch (np) {
        case 2: return launch_stage_np<2>(prec, flux, capt, P, h, s);
        case 3: return launch_stage_np<3>(prec, flux, capt, P, h, s);
        case 4: return launch_stage_np<4>(prec, flux, capt, P, h, s);
        case 5: return launch_stage_np<5>(prec, flux, capt, P, h, s);
        case 6: return launch_stage_np<6>(prec, flux, capt, P, h, s);
        default: return hipErrorInvalidValue;
    }
}

template <int NP, typename R>
static hipError_t launch_elem_t(const ElemParams& P0, const HostOps& h, int accum, int check,
                                hipStream_t s) {
    constexpr int EPB = 64;
    ElemParams P = P0;
    P.nblk = (P.nel + EPB - 1) / EPB;
    hipLaunchKernelGGL((elem_kernel<NP, R, EPB>), dim3(P.nblk), dim3(EPB * NP), 0, s, P,
                       make_ops<NP>(h), accum, check);
    return hipGetLastError();
}

hipError_t launch_elem(int np, int prec, const ElemParams& P, const HostOps& h, int accum_lambda,
                       int check_ctl, hipStream_t s) {
    switch (np) {
#define FRF_ELEM_CASE(N) \
    case N: return prec == 0 ? launch_elem_t<N, double>(P, h, accum_lambda, check_ctl, s) \
                             : launch_elem_t<N, float>(P, h, accum_lambda, check_ctl, s);
        FRF_ELEM_CASE(2)
        FRF_ELEM_CASE(3)
        FRF_ELEM_CASE(4)
        FRF_ELEM_CASE(5)
        FRF_ELEM_CASE(6)
#undef FRF_ELEM_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dt(DevCtl* c, hipStream_t s) {
    hipLaunchKernelGGL(dt_kernel, dim3(1), dim3(1), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_finish(DevCtl* c, hipStream_t s) {
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(1), 0, s, c);
    return hipGetLastError();
}

}  // namespace frf
