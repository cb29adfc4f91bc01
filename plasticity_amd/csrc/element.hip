// element.hip -- Hex8 finite-strain (F-bar) element integration for gfx950.
//
// Two kernels per assembly:
//   gp_kernel      one thread per Gauss point over all elements.  Nodal X/U of a
//                  32-element tile are staged in LDS.  Per GP: reference gradients,
//                  centroid F0/J0 (F-bar), F, Fbar = (J0/J)^(1/3) F, trial be and its
//                  eigen-decomposition, J2 radial return, P = tau Fbar^-T and the
//                  9x9 two-point tangent A = dP/dFbar, written (scaled by the
//                  quadrature weight) to a per-GP packet -- its upper triangle where A
//                  has major symmetry -- and, at the first assembly of a context, the
//                  reference gradients to the geometry array.  Restates
//                  element_force_tangent_finite! Elements.jl:405-494 with
//                  finite_kinematics FiniteStrain.jl:148-170, finite_stress_update
//                  :292-346, dPdF :503-547 and return_map Materials.jl:159-250.
//   scatter_kernel one 64-lane wave per four elements of one colour: lane (el, a, half)
//                  contracts the 8 packets of element el into the 3x3 blocks Ke(a, b),
//                  b in one half of the nodes, and adds them into the fine
//                  BSR K through a precomputed slot map (no search, no atomics:
//                  same-colour elements share no node, Assembly.jl:370-385).
//                  Ke = sum_g w G^T A Geff with Geff = s G + (s/3) vec(F)(g0 - gJ)^T
//                  (Elements.jl:470-478, :524-531).
#include <hip/hip_runtime.h>

#include "common.h"

namespace pfem {
namespace {

constexpr double EIG_TOL = 1e-9;  // FiniteStrain.jl:69

// 3x3 helpers, column-major m[i + 3j]
__device__ __forceinline__ double det3(const double *m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
           m[6] * (m[1] * m[5] - m[4] * m[2]);
}
__device__ __forceinline__ void inv3(const double *m, double *o) {
    double d = det3(m);
    o[0] = (m[4] * m[8] - m[7] * m[5]) / d;
    o[3] = -(m[3] * m[8] - m[6] * m[5]) / d;
    o[6] = (m[3] * m[7] - m[6] * m[4]) / d;
    o[1] = -(m[1] * m[8] - m[7] * m[2]) / d;
    o[4] = (m[0] * m[8] - m[6] * m[2]) / d;
    o[7] = -(m[0] * m[7] - m[6] * m[1]) / d;
    o[2] = (m[1] * m[5] - m[4] * m[2]) / d;
    o[5] = -(m[0] * m[5] - m[3] * m[2]) / d;
    o[8] = (m[0] * m[4] - m[3] * m[1]) / d;
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (Golub & Van
// Loan, Matrix Computations, Alg. 8.5.2/8.5.3).  Input: the 6 distinct
// entries; output ascending eigenvalues w and eigenvectors as columns of Q.
__device__ void eig_sym3(double a00, double a11, double a22, double a01, double a12,
                         double a02, double *w, double *Q) {
    double A[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-36 * dia || off == 0.0) break;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int p = (r == 2) ? 1 : 0;
            const int q = (r == 0) ? 1 : 2;
            double apq = A[p][q];
            if (apq == 0.0) continue;
            double theta = (A[q][q] - A[p][p]) / (2 * apq);
            double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
            double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
            A[p][q] = A[q][p] = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
    }
    double ev[3] = {A[0][0], A[1][1], A[2][2]};
    int o0 = 0, o1 = 1, o2 = 2, tmp;
    if (ev[o0] > ev[o1]) { tmp = o0; o0 = o1; o1 = tmp; }
    if (ev[o1] > ev[o2]) { tmp = o1; o1 = o2; o2 = tmp; }
    if (ev[o0] > ev[o1]) { tmp = o0; o0 = o1; o1 = tmp; }
    const int ord[3] = {o0, o1, o2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        w[c] = ev[ord[c]];
#pragma unroll
        for (int k = 0; k < 3; ++k) Q[k + 3 * c] = V[k][ord[c]];
    }
}

// dN/dxi of the trilinear hex at natural point xi (Elements.jl:62-75), node order
// of NODE_NAT (Elements.jl:26-30).
__device__ __forceinline__ void hex8_dshape(double x0, double x1, double x2, double *dN) {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const double sa = ((a + 1) & 2) ? 1.0 : -1.0;  // xi_a: -,+,+,-,-,+,+,-
        const double ta = (a & 2) ? 1.0 : -1.0;         // eta_a: -,-,+,+,-,-,+,+
        const double ua = (a & 4) ? 1.0 : -1.0;         // zeta_a
        dN[3 * a + 0] = 0.125 * sa * (1 + ta * x1) * (1 + ua * x2);
        dN[3 * a + 1] = 0.125 * ta * (1 + sa * x0) * (1 + ua * x2);
        dN[3 * a + 2] = 0.125 * ua * (1 + sa * x0) * (1 + ta * x1);
    }
}

// Reference gradients dN/dX = dN J^-1 with J = Xe^T dN (Elements.jl:89, :155-161).
__device__ __forceinline__ void ref_grads(const double *Xe, double x0, double x1, double x2,
                                          double *dNdX, double &detJ) {
    double dN[24];
    hex8_dshape(x0, x1, x2, dN);
    double J[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) s += Xe[3 * a + i] * dN[3 * a + j];
            J[i + 3 * j] = s;
        }
    detJ = det3(J);
    double Ji[9];
    inv3(J, Ji);
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            dNdX[3 * a + j] = dN[3 * a + 0] * Ji[0 + 3 * j] + dN[3 * a + 1] * Ji[1 + 3 * j] +
                              dN[3 * a + 2] * Ji[2 + 3 * j];
}

__device__ __forceinline__ void defgrad(const double *ue, const double *dNdX, double *F) {
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = (k % 4 == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) F[i + 3 * j] += ue[3 * a + i] * dNdX[3 * a + j];
}

__device__ __forceinline__ double yield_stress(const DevMat &m, double a) {
    return m.sy0 + (m.ssat - m.sy0) * (1 - exp(-m.delta * a)) + m.Hiso * a;
}
__device__ __forceinline__ double yield_slope(const DevMat &m, double a) {
    return m.delta * (m.ssat - m.sy0) * exp(-m.delta * a) + m.Hiso;
}

// Divided difference of (1/2) ln at (bA, bB) with the reference's degenerate
// switch (FiniteStrain.jl:467-476).  Off the switch, 0.5*log1p(d/bB)/d is the
// same quantity as (0.5 ln bA - 0.5 ln bB)/(bA - bB) without cancellation.
// Element k of three, by selects: the tangent's column loop indexes small arrays with
// its runtime column (p, q); selects keep those arrays in registers instead of scratch.
__device__ __forceinline__ double pick3(int k, double a0, double a1, double a2) {
    return k == 0 ? a0 : (k == 1 ? a1 : a2);
}

__device__ __forceinline__ double halflog_dd(double bA, double bB) {
    if (fabs(bA - bB) < EIG_TOL * (fabs(bA) + fabs(bB) + 1.0)) return 1.0 / (bA + bB);
    double d = bA - bB;
    return 0.5 * log1p(d / bB) / d;
}

// KIN: some material has Hkin > 0.  Those Gauss points push the stored back stress
// forward with the polar rotation of Fbar and add the back-stress rotation term to A
// (FiniteStrain.jl:190-262, :312-320, :533-537); the KIN = false instantiation is the
// isotropic kernel without that code.
template <bool KIN>
__global__ __launch_bounds__(256) void gp_kernel(ElementArgs a, int64_t e_begin, int64_t e_end) {
    __shared__ double sX[32][24];
    __shared__ double sU[32][24];
    const int le = threadIdx.x >> 3, g = threadIdx.x & 7;
    const int64_t li = e_begin + (int64_t)blockIdx.x * 32 + le;  // list position
    const bool valid = li < e_end;
    const int64_t e = valid ? (a.elist ? a.elist[li] : li) : 0;
    if (valid) {
        const int64_t node = a.conn[8 * e + g];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sX[le][3 * g + d] = a.X[3 * node + d];
            sU[le][3 * g + d] = a.U[3 * node + d];
        }
    }
    __syncthreads();
    if (!valid) return;
    const double *Xe = sX[le];
    const double *ue = sU[le];
    const DevMat m = a.mats[a.matid ? a.matid[e] : 0];
    const int64_t gp = 8 * e + g;

    // --- geometry at this GP (Elements.jl:130-161): GAUSS_PTS order :34-42
    const double gq = 0.57735026918962576451;  // 1/sqrt(3)
    const double x0 = (((g + 1) & 2) ? gq : -gq), x1 = ((g & 2) ? gq : -gq), x2 = ((g & 4) ? gq : -gq);
    double dNdX[24], detJ;
    ref_grads(Xe, x0, x1, x2, dNdX, detJ);
    const double w = detJ;  // GAUSS_WTS == 1
    if (a.geo_write)
#pragma unroll
        for (int k = 0; k < 24; ++k) a.geo[24 * gp + k] = dNdX[k];

    // --- F-bar centroid quantities (Elements.jl:421-424, :517-519)
    double J0 = 1.0, F0i[9];
    if (a.fbar) {
        double dNdX0[24], dj0, F0[9];
        ref_grads(Xe, 0.0, 0.0, 0.0, dNdX0, dj0);
        defgrad(ue, dNdX0, F0);
        J0 = det3(F0);
        inv3(F0, F0i);
        if (g == 0) {  // g0 = G0^T vec(F0^-T) (Elements.jl:528)
#pragma unroll
            for (int n = 0; n < 8; ++n)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    a.g0[24 * e + 3 * n + k] = dNdX0[3 * n + 0] * F0i[0 + 3 * k] +
                                               dNdX0[3 * n + 1] * F0i[1 + 3 * k] +
                                               dNdX0[3 * n + 2] * F0i[2 + 3 * k];
        }
    }
    double F[9];
    defgrad(ue, dNdX, F);
    const double Jg = det3(F);
    const double s = a.fbar ? cbrt(J0 / Jg) : 1.0;  // Elements.jl:437
    double Fb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fb[k] = s * F[k];
    const double Jb = det3(Fb);
    if (Jb <= 0) {  // ElementInversionError (FiniteStrain.jl:151, Elements.jl:447)
        atomicMin(a.fail_key, (unsigned long long)gp);
        return;
    }

    // --- trial elastic left Cauchy-Green be = Fb Cpi Fb^T (FiniteStrain.jl:155-156)
    const double *cp = a.cpinv + 6 * gp;
    const double c0 = cp[0], c1 = cp[1], c2 = cp[2], c3 = cp[3], c4 = cp[4], c5 = cp[5];
    double M[9];  // M = Fb * Cpi (column-major)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        M[i + 0] = Fb[i] * c0 + Fb[i + 3] * c3 + Fb[i + 6] * c5;
        M[i + 3] = Fb[i] * c3 + Fb[i + 3] * c1 + Fb[i + 6] * c4;
        M[i + 6] = Fb[i] * c5 + Fb[i + 3] * c4 + Fb[i + 6] * c2;
    }
    double be[6];  // (00,11,22,01,12,02)
    {
        const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int t = 0; t < 6; ++t)
            be[t] = M[I[t]] * Fb[Jx[t]] + M[I[t] + 3] * Fb[Jx[t] + 3] + M[I[t] + 6] * Fb[Jx[t] + 6];
    }
    double b[3], Q[9];
    eig_sym3(be[0], be[1], be[2], be[3], be[4], be[5], b, Q);
    const double l0 = 0.5 * log(b[0]), l1 = 0.5 * log(b[1]), l2 = 0.5 * log(b[2]);
    double ee[6];  // trial Hencky strain, engineering shear (FiniteStrain.jl:162-168)
    {
        const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            double v = l0 * Q[I[t]] * Q[Jx[t]] + l1 * Q[I[t] + 3] * Q[Jx[t] + 3] +
                       l2 * Q[I[t] + 6] * Q[Jx[t] + 6];
            ee[t] = (t < 3) ? v : 2 * v;
        }
    }

    // --- objective back stress (FiniteStrain.jl:312-314): Fb = R U, beta_sp = R beta_ref R^T.
    //     Kept in U's eigenbasis V: RV = R V = Fb V diag(1/sqrt(lp)), bv = V^T beta_ref V.
    const bool kin = KIN && m.Hkin > 0;
    double RV[9], Vp[9], lp[3], bv[6], bsp[6] = {0, 0, 0, 0, 0, 0};
    if (kin) {
        double cc[6];  // C = Fb^T Fb (00,11,22,01,12,02)
        {
            const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int t = 0; t < 6; ++t)
                cc[t] = Fb[3 * I[t]] * Fb[3 * Jx[t]] + Fb[3 * I[t] + 1] * Fb[3 * Jx[t] + 1] +
                        Fb[3 * I[t] + 2] * Fb[3 * Jx[t] + 2];
        }
        double wc[3];
        eig_sym3(cc[0], cc[1], cc[2], cc[3], cc[4], cc[5], wc, Vp);
#pragma unroll
        for (int A_ = 0; A_ < 3; ++A_) {
            lp[A_] = sqrt(wc[A_] > 0.0 ? wc[A_] : 0.0);  // FiniteStrain.jl:195-197
            const double is = lp[A_] > 0 ? 1.0 / lp[A_] : 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
                RV[i + 3 * A_] = is * (Fb[i] * Vp[3 * A_] + Fb[i + 3] * Vp[3 * A_ + 1] + Fb[i + 6] * Vp[3 * A_ + 2]);
        }
        const double *br = a.beta + 6 * gp;
        const double B[9] = {br[0], br[3], br[5], br[3], br[1], br[4], br[5], br[4], br[2]};
        double BV[9];  // B V
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int A_ = 0; A_ < 3; ++A_)
                BV[i + 3 * A_] = B[i] * Vp[3 * A_] + B[i + 3] * Vp[3 * A_ + 1] + B[i + 6] * Vp[3 * A_ + 2];
        {
            const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int t = 0; t < 6; ++t)
                bv[t] = Vp[3 * I[t]] * BV[3 * Jx[t]] + Vp[3 * I[t] + 1] * BV[3 * Jx[t] + 1] +
                        Vp[3 * I[t] + 2] * BV[3 * Jx[t] + 2];
            const double Bv[9] = {bv[0], bv[3], bv[5], bv[3], bv[1], bv[4], bv[5], bv[4], bv[2]};
            double T1[9];  // RV Bv
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int A_ = 0; A_ < 3; ++A_)
                    T1[i + 3 * A_] = RV[i] * Bv[3 * A_] + RV[i + 3] * Bv[3 * A_ + 1] + RV[i + 6] * Bv[3 * A_ + 2];
#pragma unroll
            for (int t = 0; t < 6; ++t)
                bsp[t] = T1[I[t]] * RV[Jx[t]] + T1[I[t] + 3] * RV[Jx[t] + 3] + T1[I[t] + 6] * RV[Jx[t] + 6];
        }
    }

    // --- J2 radial return with zero additive plastic strain and the spatial back stress
    //     (Materials.jl:159-250 as called at FiniteStrain.jl:318)
    const double G = m.G, lam = m.lam;
    const double tr = ee[0] + ee[1] + ee[2];
    double tau[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) tau[k] = lam * tr + 2 * G * ee[k];
#pragma unroll
    for (int k = 3; k < 6; ++k) tau[k] = G * ee[k];
    const double p = (tau[0] + tau[1] + tau[2]) / 3;
    double xi[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) xi[k] = tau[k] - (k < 3 ? p : 0.0) - bsp[k];
    const double nrm = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                            2 * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
    const double s32 = 1.2247448713915890491;  // sqrt(3/2)
    const double qtr = s32 * nrm;
    const double an = a.abar[gp];
    const double ftr = qtr - yield_stress(m, an);
    double cdev = 0.0, cnn = 0.0, dg = 0.0;  // D = C - cdev*Idev + cnn*nh nh^T
    double nh[6] = {0, 0, 0, 0, 0, 0};
    const bool plastic = ftr > 0.0;
    if (plastic) {
#pragma unroll
        for (int k = 0; k < 6; ++k) nh[k] = xi[k] / nrm;
        if (m.delta > 0 && m.ssat > m.sy0) {  // Voce local Newton (Materials.jl:201-213)
            dg = ftr / (3 * G + m.Hkin + yield_slope(m, an));
            for (int it = 0; it < 50; ++it) {
                double Hp = yield_slope(m, an + dg);
                double r = qtr - (3 * G + m.Hkin) * dg - yield_stress(m, an + dg);
                double ddg = r / (3 * G + m.Hkin + Hp);
                dg += ddg;
                if (fabs(ddg) <= 1e-13 * (dg + 1e-14)) break;
            }
        } else {
            dg = ftr / (3 * G + m.Hiso + m.Hkin);  // Materials.jl:218
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) tau[k] -= (2 * G * dg * s32) * nh[k];
        const double Hp = yield_slope(m, an + dg);
        const double b0 = 2 * G * dg / qtr, g0c = 2 * G / (3 * G + m.Hkin + Hp);
        cdev = 3 * G * b0;
        cnn = 3 * G * (b0 - g0c);
    }
    // --- P = tau Fb^-T (FiniteStrain.jl:560-561)
    double Fbi[9];
    inv3(Fb, Fbi);
    double T[9] = {tau[0], tau[3], tau[5], tau[3], tau[1], tau[4], tau[5], tau[4], tau[2]};
    double Fit[9];  // Fb^-T
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Fit[r + 3 * c] = Fbi[c + 3 * r];
    double P[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            P[i + 3 * j] = T[i] * Fit[0 + 3 * j] + T[i + 3] * Fit[1 + 3 * j] + T[i + 6] * Fit[2 + 3 * j];

    // --- Kirchhoff stress out (Elements.jl:481-482)
    double *sg = a.sig + 6 * gp;
#pragma unroll
    for (int k = 0; k < 6; ++k) sg[k] = tau[k];

    if (a.commit) {  // exp-map plastic update (FiniteStrain.jl:329-343), Elements.jl:483-491
        double dep[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) dep[k] = (dg * s32) * (k < 3 ? nh[k] : 2 * nh[k]);
        double ev[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) ev[k] = ee[k] - dep[k];
        double wv[3], V[9];
        eig_sym3(2 * ev[0], 2 * ev[1], 2 * ev[2], ev[3], ev[4], ev[5], wv, V);
        const double x0e = exp(wv[0]), x1e = exp(wv[1]), x2e = exp(wv[2]);
        double Bn[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Bn[i + 3 * j] = x0e * V[i] * V[j] + x1e * V[i + 3] * V[j + 3] + x2e * V[i + 6] * V[j + 6];
        double Tm[9];  // Fbi * Bn
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Tm[i + 3 * j] = Fbi[i] * Bn[0 + 3 * j] + Fbi[i + 3] * Bn[1 + 3 * j] + Fbi[i + 6] * Bn[2 + 3 * j];
        double cn[6];
        {
            const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int t = 0; t < 6; ++t)
                cn[t] = Tm[I[t]] * Fbi[Jx[t]] + Tm[I[t] + 3] * Fbi[Jx[t] + 3] + Tm[I[t] + 6] * Fbi[Jx[t] + 6];
        }
        double *ep = a.epsp + 6 * gp;
        double *cw = a.cpinv + 6 * gp;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            ep[k] += dep[k];
            cw[k] = cn[k];
        }
        a.abar[gp] = an + dg;
        if (KIN) {  // beta_ref_new = R^T beta_sp_new R (FiniteStrain.jl:230, :320); 0 when Hkin == 0
            double *bw = a.beta + 6 * gp;
            if (kin) {
                const double c = (2.0 / 3.0) * m.Hkin * (dg * s32);
                double bn[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) bn[k] = bsp[k] + c * nh[k];
                const double Bn[9] = {bn[0], bn[3], bn[5], bn[3], bn[1], bn[4], bn[5], bn[4], bn[2]};
                double Rm[9];  // R = RV V^T
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        Rm[i + 3 * j] = RV[i] * Vp[j] + RV[i + 3] * Vp[j + 3] + RV[i + 6] * Vp[j + 6];
                double T1[9];  // Bn R
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        T1[i + 3 * j] = Bn[i] * Rm[3 * j] + Bn[i + 3] * Rm[3 * j + 1] + Bn[i + 6] * Rm[3 * j + 2];
                const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
                for (int t = 0; t < 6; ++t)
                    bw[t] = Rm[3 * I[t]] * T1[3 * Jx[t]] + Rm[3 * I[t] + 1] * T1[3 * Jx[t] + 1] +
                            Rm[3 * I[t] + 2] * T1[3 * Jx[t] + 2];
            } else {
#pragma unroll
                for (int k = 0; k < 6; ++k) bw[k] = 0.0;
            }
        }
    }
    if (!a.packet) return;  // state-only pass (the commit of a converged load step)

    // dtau/dbeta_sp = ka I + kc xi (W xi)^T in the plastic branch (FiniteStrain.jl:245-262)
    double ka = 0.0, kc = 0.0;
    if (kin && plastic) {
        const double Hs = m.Hiso + m.Hkin, Y = m.sy0 + m.Hiso * an;
        const double dgk = (s32 * nrm - Y) / (3 * G + Hs);
        ka = 2 * G * s32 * dgk / nrm;
        kc = (2 * G * s32 / (3 * G + Hs) * (Y / (nrm * nrm))) / nrm;
    }

    constexpr bool SYM = !KIN;  // A has major symmetry without the back-stress rotation term
    constexpr int NA = pk_na(SYM);
    double *pk = a.packet + (size_t)(8 * (li - e_begin) + g) * pk_len(SYM);
    const double sw = s * w;

    // --- A = dP/dFbar, one column per unit perturbation dF = e_p e_q^T
    //     (FiniteStrain.jl:503-547).  dbe = e_p m^T + m e_p^T with m = (Fb Cpi)[:,q].
    double gam[3][3];
#pragma unroll
    for (int A_ = 0; A_ < 3; ++A_)
#pragma unroll
        for (int B_ = 0; B_ < 3; ++B_)
            gam[A_][B_] = (A_ == B_) ? 1.0 / (2 * b[A_]) : halflog_dd(b[A_], b[B_]);
    double qv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int col = 0; col < 9; ++col) {
        const int pp = col % 3, qq = col / 3;
        // column qq of M and of Fb^-T, row pp of Fb^-T, F[col]
        double mq[3], fq[3], fp[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            mq[i] = pick3(qq, M[i], M[i + 3], M[i + 6]);
            fq[i] = pick3(qq, Fit[i], Fit[i + 3], Fit[i + 6]);
            fp[i] = pick3(pp, Fit[3 * i], Fit[3 * i + 1], Fit[3 * i + 2]);
        }
        const double fcol = pick3(qq, pick3(pp, F[0], F[1], F[2]), pick3(pp, F[3], F[4], F[5]),
                                  pick3(pp, F[6], F[7], F[8]));
        double u[3], v[3];
#pragma unroll
        for (int A_ = 0; A_ < 3; ++A_) {
            u[A_] = pick3(pp, Q[3 * A_], Q[1 + 3 * A_], Q[2 + 3 * A_]);
            v[A_] = Q[0 + 3 * A_] * mq[0] + Q[1 + 3 * A_] * mq[1] + Q[2 + 3 * A_] * mq[2];
        }
        double Y[3][3];  // gamma o (Q^T dbe Q)
#pragma unroll
        for (int A_ = 0; A_ < 3; ++A_)
#pragma unroll
            for (int B_ = 0; B_ < 3; ++B_) Y[A_][B_] = gam[A_][B_] * (u[A_] * v[B_] + v[A_] * u[B_]);
        double QY[9];  // Q * Y
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int B_ = 0; B_ < 3; ++B_)
                QY[i + 3 * B_] = Q[i] * Y[0][B_] + Q[i + 3] * Y[1][B_] + Q[i + 6] * Y[2][B_];
        double de[6];  // d eps_e, engineering shear (FiniteStrain.jl:485)
        {
            const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                double vv = QY[I[t]] * Q[Jx[t]] + QY[I[t] + 3] * Q[Jx[t] + 3] + QY[I[t] + 6] * Q[Jx[t] + 6];
                de[t] = (t < 3) ? vv : 2 * vv;
            }
        }
        // d tau = D de with D = C - cdev*Idev + cnn*nh nh^T (Materials.jl:247)
        const double trd = de[0] + de[1] + de[2];
        double nde = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) nde += nh[k] * de[k];
        double dt[6];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            dt[k] = lam * trd + 2 * G * de[k] - cdev * (de[k] - trd / 3) + cnn * nde * nh[k];
#pragma unroll
        for (int k = 3; k < 6; ++k) dt[k] = G * de[k] - cdev * 0.5 * de[k] + cnn * nde * nh[k];
        if (kin && plastic) {  // + dtau/dbeta_sp : d beta_sp (FiniteStrain.jl:533-537)
            // Omega_b = V^T R^T dR skew: (x y^T - y x^T)_AB / (lp_A + lp_B), x = (RV)^T e_p, y = V^T e_q
            double x[3], y[3];
#pragma unroll
            for (int A_ = 0; A_ < 3; ++A_) {
                x[A_] = pick3(pp, RV[3 * A_], RV[1 + 3 * A_], RV[2 + 3 * A_]);
                y[A_] = pick3(qq, Vp[3 * A_], Vp[1 + 3 * A_], Vp[2 + 3 * A_]);
            }
            double Om[3][3];
#pragma unroll
            for (int A_ = 0; A_ < 3; ++A_)
#pragma unroll
                for (int B_ = 0; B_ < 3; ++B_)
                    Om[A_][B_] = (A_ == B_) ? 0.0 : (x[A_] * y[B_] - y[A_] * x[B_]) / (lp[A_] + lp[B_]);
            const double Bv[3][3] = {{bv[0], bv[3], bv[5]}, {bv[3], bv[1], bv[4]}, {bv[5], bv[4], bv[2]}};
            double OB[3][3];
#pragma unroll
            for (int A_ = 0; A_ < 3; ++A_)
#pragma unroll
                for (int B_ = 0; B_ < 3; ++B_)
                    OB[A_][B_] = Om[A_][0] * Bv[0][B_] + Om[A_][1] * Bv[1][B_] + Om[A_][2] * Bv[2][B_];
            double Mb[3][3];  // Om Bv - Bv Om = OB + OB^T (Om skew, Bv symmetric)
#pragma unroll
            for (int A_ = 0; A_ < 3; ++A_)
#pragma unroll
                for (int B_ = 0; B_ < 3; ++B_) Mb[A_][B_] = OB[A_][B_] + OB[B_][A_];
            double T1[9];  // RV Mb
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int B_ = 0; B_ < 3; ++B_)
                    T1[i + 3 * B_] = RV[i] * Mb[0][B_] + RV[i + 3] * Mb[1][B_] + RV[i + 6] * Mb[2][B_];
            double db[6];  // d beta_sp = RV Mb RV^T, physical-shear Voigt
            {
                const int I[6] = {0, 1, 2, 0, 1, 0}, Jx[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
                for (int t = 0; t < 6; ++t)
                    db[t] = T1[I[t]] * RV[Jx[t]] + T1[I[t] + 3] * RV[Jx[t] + 3] + T1[I[t] + 6] * RV[Jx[t] + 6];
            }
            const double xw = xi[0] * db[0] + xi[1] * db[1] + xi[2] * db[2] +
                              2 * (xi[3] * db[3] + xi[4] * db[4] + xi[5] * db[5]);
#pragma unroll
            for (int k = 0; k < 6; ++k) dt[k] += ka * db[k] + kc * xi[k] * xw;
        }
        const double dT[9] = {dt[0], dt[3], dt[5], dt[3], dt[1], dt[4], dt[5], dt[4], dt[2]};
        // dP = dtau Fb^-T - (tau Fb^-T[:,q]) (Fb^-T[p,:])   (FiniteStrain.jl:539-540)
        double wq[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) wq[r] = T[r] * fq[0] + T[r + 3] * fq[1] + T[r + 6] * fq[2];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double dP = dT[r] * Fit[0 + 3 * c] + dT[r + 3] * Fit[1 + 3 * c] + dT[r + 6] * Fit[2 + 3 * c] -
                            wq[r] * fp[c];
                if (SYM) {
                    if (r + 3 * c <= col) pk[col * (col + 1) / 2 + r + 3 * c] = sw * dP;
                } else {
                    pk[(r + 3 * c) + 9 * col] = sw * dP;
                }
                qv[r + 3 * c] += dP * fcol;
            }
    }
    if (a.fbar) {
        double Fi[9];
        inv3(F, Fi);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            pk[NA + k] = (sw / 3) * qv[k];
            pk[NA + 9 + k] = Fi[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) pk[NA + 18 + k] = w * P[k];

}

// Small-strain Hex8 (element_force_tangent! Elements.jl:323-360): eps = B ue with the
// reference-configuration gradients (element_geometry :130-144, bmatrix :97-118), the
// full return map with the committed eps_p, beta, abar (Materials.jl:159-250), sigma to
// the stress output.  The packet carries the same two-point form as the finite-strain
// kernel so that the scatter is shared: with H = grad u, eps_v(ij) = H_ij + H_ji off the
// diagonal, so  B^T D B = G^T A G  with  A_(i+3j),(k+3l) = D_v(ij),v(kl)  and
// B^T sigma = G^T vec(sigma).
__global__ __launch_bounds__(256) void gp_small_kernel(ElementArgs a, int64_t e_begin, int64_t e_end) {
    __shared__ double sX[32][24];
    __shared__ double sU[32][24];
    const int le = threadIdx.x >> 3, g = threadIdx.x & 7;
    const int64_t li = e_begin + (int64_t)blockIdx.x * 32 + le;  // list position
    const bool valid = li < e_end;
    const int64_t e = valid ? (a.elist ? a.elist[li] : li) : 0;
    if (valid) {
        const int64_t node = a.conn[8 * e + g];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sX[le][3 * g + d] = a.X[3 * node + d];
            sU[le][3 * g + d] = a.U[3 * node + d];
        }
    }
    __syncthreads();
    if (!valid) return;
    const double *Xe = sX[le];
    const double *ue = sU[le];
    const DevMat m = a.mats[a.matid ? a.matid[e] : 0];
    const int64_t gp = 8 * e + g;
    const double gq = 0.57735026918962576451;
    const double x0 = (((g + 1) & 2) ? gq : -gq), x1 = ((g & 2) ? gq : -gq), x2 = ((g & 4) ? gq : -gq);
    double dNdX[24], w;
    ref_grads(Xe, x0, x1, x2, dNdX, w);  // w = detJ (GAUSS_WTS == 1)
    if (a.geo_write)
#pragma unroll
        for (int k = 0; k < 24; ++k) a.geo[24 * gp + k] = dNdX[k];
    double H[9];
    defgrad(ue, dNdX, H);
    H[0] -= 1.0; H[4] -= 1.0; H[8] -= 1.0;
    const double eps[6] = {H[0], H[4], H[8], H[3] + H[1], H[7] + H[5], H[6] + H[2]};

    // --- return_map (Materials.jl:159-250)
    const double G = m.G, lam = m.lam;
    const double *epn = a.epsp + 6 * gp;
    const double *bn = a.beta + 6 * gp;
    double d[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = eps[k] - epn[k];
    const double tr = d[0] + d[1] + d[2];
    double sg[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) sg[k] = lam * tr + 2 * G * d[k];
#pragma unroll
    for (int k = 3; k < 6; ++k) sg[k] = G * d[k];
    const double p = (sg[0] + sg[1] + sg[2]) / 3;
    double xi[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) xi[k] = sg[k] - (k < 3 ? p : 0.0) - bn[k];
    const double nrm = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                            2 * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
    const double s32 = 1.2247448713915890491;  // sqrt(3/2)
    const double qtr = s32 * nrm;
    const double an = a.abar[gp];
    const double ftr = qtr - yield_stress(m, an);
    double cdev = 0.0, cnn = 0.0, dg = 0.0;  // D = C - cdev*Idev + cnn*nh nh^T
    double nh[6] = {0, 0, 0, 0, 0, 0};
    const bool plastic = ftr > 0.0;
    if (plastic) {
#pragma unroll
        for (int k = 0; k < 6; ++k) nh[k] = xi[k] / nrm;
        if (m.delta > 0 && m.ssat > m.sy0) {  // Voce local Newton (Materials.jl:201-213)
            dg = ftr / (3 * G + m.Hkin + yield_slope(m, an));
            for (int it = 0; it < 50; ++it) {
                double Hp = yield_slope(m, an + dg);
                double r = qtr - (3 * G + m.Hkin) * dg - yield_stress(m, an + dg);
                double ddg = r / (3 * G + m.Hkin + Hp);
                dg += ddg;
                if (fabs(ddg) <= 1e-13 * (dg + 1e-14)) break;
            }
        } else {
            dg = ftr / (3 * G + m.Hiso + m.Hkin);  // Materials.jl:218
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) sg[k] -= (2 * G * dg * s32) * nh[k];
        const double Hp = yield_slope(m, an + dg);
        const double b0 = 2 * G * dg / qtr, g0c = 2 * G / (3 * G + m.Hkin + Hp);
        cdev = 3 * G * b0;
        cnn = 3 * G * (b0 - g0c);
    }
    double *so = a.sig + 6 * gp;
#pragma unroll
    for (int k = 0; k < 6; ++k) so[k] = sg[k];
    if (a.commit) {  // Elements.jl:350-357
        double *ep = a.epsp + 6 * gp;
        double *bw = a.beta + 6 * gp;
        const double c = (2.0 / 3.0) * m.Hkin * (dg * s32);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            ep[k] += (dg * s32) * (k < 3 ? nh[k] : 2 * nh[k]);
            bw[k] += c * nh[k];
        }
        a.abar[gp] = an + dg;
    }
    if (!a.packet) return;

    // D (6x6) = C - cdev Idev + cnn nh nh^T, scattered to A_(i+3j),(k+3l) = w D_v(ij),v(kl)
    double Dm[6][6];
#pragma unroll
    for (int v = 0; v < 6; ++v)
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            const double Cvu = (v < 3 && u < 3) ? (lam + (v == u ? 2 * G : 0.0)) : (v == u ? G : 0.0);
            const double Idev = (v == u ? (v < 3 ? 1.0 : 0.5) : 0.0) - ((v < 3 && u < 3) ? 1.0 / 3.0 : 0.0);
            Dm[v][u] = w * (Cvu - cdev * Idev + cnn * nh[v] * nh[u]);
        }
    double *pk = a.packet + (size_t)(8 * (li - e_begin) + g) * PK_SYM;  // D symmetric: packed layout
    const int VO[9] = {0, 3, 5, 3, 1, 4, 5, 4, 2};  // Voigt index of (i, j), i + 3j
#pragma unroll
    for (int c = 0; c < 9; ++c)
#pragma unroll
        for (int r = 0; r <= c; ++r) pk[c * (c + 1) / 2 + r] = Dm[VO[r]][VO[c]];
#pragma unroll
    for (int k = 0; k < 9; ++k) pk[pk_na(true) + 18 + k] = w * sg[VO[k]];
}

// Entry (p, q) = (i + 3j, k + 3l) of s w A in a Gauss-point packet of either layout.
template <bool SYM>
__device__ __forceinline__ int pk_a_index(int p, int q) {
    return SYM ? pk_sym_index(p, q) : p + 9 * q;
}

// Contraction of the 8 GP packets of an element into 3x3 blocks Ke(a, b),
// b in [4*half, 4*half+4), plus the residual part of node a (half == 0).  ge: the
// element's 8 x 24 reference gradients.
template <bool SYM>
__device__ __forceinline__ void contract_blocks(const double *__restrict__ pe, const double *__restrict__ ge,
                                                const double *__restrict__ g0e, int a, int half, int fbar,
                                                double acc[4][9], double fe[3]) {
    constexpr int LEN = pk_len(SYM), NA = pk_na(SYM);
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[bb][t] = 0.0;
    fe[0] = fe[1] = fe[2] = 0.0;
    double g0b[4][3];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
        for (int k = 0; k < 3; ++k) g0b[bb][k] = fbar ? g0e[3 * (4 * half + bb) + k] : 0.0;
#pragma unroll 1
    for (int g = 0; g < 8; ++g) {
        const double *pk = pe + g * LEN;
        const double *dn = ge + 24 * g;
        const double dNa0 = dn[3 * a], dNa1 = dn[3 * a + 1], dNa2 = dn[3 * a + 2];
        double Ua[3][9];  // Ua[i][(k,l)] = sum_j dNa_j SA[(i,j),(k,l)]
#pragma unroll
        for (int c = 0; c < 9; ++c)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                Ua[i][c] = dNa0 * pk[pk_a_index<SYM>(i, c)] + dNa1 * pk[pk_a_index<SYM>(i + 3, c)] +
                           dNa2 * pk[pk_a_index<SYM>(i + 6, c)];
        double h[3] = {0, 0, 0}, Fi[9];
        if (fbar) {
#pragma unroll
            for (int i = 0; i < 3; ++i) h[i] = dNa0 * pk[NA + i] + dNa1 * pk[NA + 3 + i] + dNa2 * pk[NA + 6 + i];
#pragma unroll
            for (int k = 0; k < 9; ++k) Fi[k] = pk[NA + 9 + k];
        }
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int bn = 4 * half + bb;
            const double d0 = dn[3 * bn], d1 = dn[3 * bn + 1], d2 = dn[3 * bn + 2];
            double gd[3] = {0, 0, 0};
            if (fbar) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    gd[k] = g0b[bb][k] - (d0 * Fi[0 + 3 * k] + d1 * Fi[1 + 3 * k] + d2 * Fi[2 + 3 * k]);
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    acc[bb][3 * i + k] += Ua[i][k] * d0 + Ua[i][k + 3] * d1 + Ua[i][k + 6] * d2 + h[i] * gd[k];
        }
        if (half == 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                fe[i] += dNa0 * pk[NA + 18 + i] + dNa1 * pk[NA + 21 + i] + dNa2 * pk[NA + 24 + i];
        }
    }
}

// One 64-lane wave per four elements of one colour; lane (el, a, half) =
// (lane / 16, lane % 16 / 2, lane % 2) accumulates the four 3x3 blocks Ke(a, b),
// b = 4 half .. 4 half + 3, of element el over the 8 GP packets and adds them into the
// fine BSR K through the precomputed slot map (same-colour elements share no node, so
// no atomics; colours run serially, Assembly.jl:370-385).
//   Ke(a,b)_ik = sum_g [ sum_jl dNa_j dNb_l SA_(i+3j),(k+3l) - h_i gJ_b,k ] + g0_b,k sum_g h_i
//   h_i = sum_j dNa_j q_(i+3j),  gJ_b,k = sum_j dNb_j (F^-1)_jk        (Elements.jl:470-478)
// Per Gauss point the lane runs over the 9 columns c = k + 3l of SA: it forms
// U_i = sum_j dNa_j SA_(i+3j),c and adds U_i dNb_l to its four blocks, so every SA entry
// is read from LDS once per lane and the 16 lanes of an element read 4x less than one
// lane per block did (LDS bandwidth was the limit).  Each Gauss point's packets and
// reference gradients (four elements) are staged in a double buffer, the next point's
// 16-byte loads in flight during the contraction.
// Gauss-point packets and reference gradients are read once per assembly: non-temporal
// loads keep them from evicting the K blocks the colour passes revisit.
typedef double d2v_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 stream_load2(const double2 *p) {
    const d2v_t v = __builtin_nontemporal_load(reinterpret_cast<const d2v_t *>(p));
    return double2{v.x, v.y};
}

template <bool SYM>
__global__ __launch_bounds__(64) void scatter_kernel(const double *__restrict__ packet,
                                                      const double *__restrict__ geo,
                                                      const double *__restrict__ g0,
                                                      const int64_t *__restrict__ elist, int64_t count,
                                                      const int64_t *__restrict__ conn,
                                                      const int32_t *__restrict__ slots,
                                                      const uint64_t *__restrict__ first,
                                                      double *__restrict__ Kv, double *__restrict__ R, int fbar,
                                                      int local) {
    constexpr int LEN = pk_len(SYM), NA = pk_na(SYM);
    constexpr int ST = LEN + 24;               // doubles staged per element and Gauss point
    constexpr int NCP = LEN / 2, NC = ST / 2;  // 16-byte chunks: packet, then the 24 gradients
    constexpr int NCH = (4 * NC + 63) / 64;    // chunks per lane and Gauss point
    __shared__ __attribute__((aligned(16))) double sp[2][4 * ST];
    const int lane = threadIdx.x, el = lane >> 4, a = (lane >> 1) & 7, half = lane & 1;
    const int64_t li0 = (int64_t)blockIdx.x * 4;
    const bool act = li0 + el < count;
    const int64_t e = act ? elist[li0 + el] : 0;
    // the lane's chunks: source of Gauss point 0 and the stride per Gauss point (double2)
    const double2 *csrc[NCH];
    int cstride[NCH];
    bool cok[NCH];
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        const int q = lane + 64 * k, ce = q / NC, t = q - ce * NC;
        cok[k] = q < 4 * NC && li0 + ce < count;
        const int64_t cli = cok[k] ? li0 + ce : 0;
        const int64_t ee = cok[k] ? elist[cli] : 0;
        if (t < NCP) {
            csrc[k] = reinterpret_cast<const double2 *>(packet + (size_t)(local ? cli : ee) * 8 * LEN) + t;
            cstride[k] = NCP;
        } else {
            csrc[k] = reinterpret_cast<const double2 *>(geo + (size_t)ee * 8 * 24) + (t - NCP);
            cstride[k] = 12;
        }
    }
    double2 rb[NCH];
#pragma unroll
    for (int k = 0; k < NCH; ++k) rb[k] = cok[k] ? stream_load2(csrc[k]) : double2{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NCH; ++k)
        if (lane + 64 * k < 4 * NC) reinterpret_cast<double2 *>(sp[0])[lane + 64 * k] = rb[k];
    __syncthreads();
    double acc[4][9];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[bb][t] = 0.0;
    double fe[3] = {0.0, 0.0, 0.0}, H[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
    for (int g = 0; g < 8; ++g) {
        if (g < 7) {  // the next Gauss point, written to the other buffer at the end
#pragma unroll
            for (int k = 0; k < NCH; ++k)
                if (cok[k]) rb[k] = stream_load2(csrc[k] + (g + 1) * cstride[k]);
        }
        const double *pk = sp[g & 1] + el * ST;
        double dNa[3], dNb[4][3];
#pragma unroll
        for (int j = 0; j < 3; ++j) dNa[j] = pk[LEN + 3 * a + j];
#pragma unroll
        for (int bb = 0; bb < 4; ++bb)
#pragma unroll
            for (int j = 0; j < 3; ++j) dNb[bb][j] = pk[LEN + 3 * (4 * half + bb) + j];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const int k = c % 3, l = c / 3;
            double u[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                u[i] = dNa[0] * pk[pk_a_index<SYM>(i, c)] + dNa[1] * pk[pk_a_index<SYM>(i + 3, c)] +
                       dNa[2] * pk[pk_a_index<SYM>(i + 6, c)];
#pragma unroll
            for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                for (int i = 0; i < 3; ++i) acc[bb][3 * i + k] += u[i] * dNb[bb][l];
        }
        if (fbar) {
            double h[3], Fi[9];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                h[i] = dNa[0] * pk[NA + i] + dNa[1] * pk[NA + 3 + i] + dNa[2] * pk[NA + 6 + i];
                H[i] += h[i];
            }
#pragma unroll
            for (int t = 0; t < 9; ++t) Fi[t] = pk[NA + 9 + t];
#pragma unroll
            for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double gJ = dNb[bb][0] * Fi[3 * k] + dNb[bb][1] * Fi[3 * k + 1] + dNb[bb][2] * Fi[3 * k + 2];
#pragma unroll
                    for (int i = 0; i < 3; ++i) acc[bb][3 * i + k] -= h[i] * gJ;
                }
        }
        if (half == 0)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                fe[i] += dNa[0] * pk[NA + 18 + i] + dNa[1] * pk[NA + 21 + i] + dNa[2] * pk[NA + 24 + i];
        if (g < 7) {
#pragma unroll
            for (int k = 0; k < NCH; ++k)
                if (lane + 64 * k < 4 * NC) reinterpret_cast<double2 *>(sp[(g + 1) & 1])[lane + 64 * k] = rb[k];
        }
        __syncthreads();
    }
    if (!act) return;
    if (fbar)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double g0k = g0[24 * e + 3 * (4 * half + bb) + k];
#pragma unroll
                for (int i = 0; i < 3; ++i) acc[bb][3 * i + k] += g0k * H[i];
            }
    // the block's first contribution in colour order is stored (K is not cleared
    // before the assembly), every later one added
    const uint64_t fm = first[e];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) {
        const int b = 4 * half + bb;
        double *kb = Kv + 9 * (int64_t)slots[64 * e + 8 * a + b];
        if ((fm >> (8 * a + b)) & 1) {
#pragma unroll
            for (int t = 0; t < 9; ++t) kb[t] = acc[bb][t];
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) kb[t] += acc[bb][t];
        }
    }
    if (half == 0) {
        const int64_t n = conn[8 * e + a];
#pragma unroll
        for (int i = 0; i < 3; ++i) R[3 * n + i] += fe[i];
    }
}

template <bool SYM>
__global__ __launch_bounds__(256) void dense_ke_kernel(const double *__restrict__ packet,
                                                       const double *__restrict__ geo,
                                                       const double *__restrict__ g0, int64_t nelem,
                                                       double *__restrict__ Fe, double *__restrict__ Ke,
                                                       int fbar) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t e = t >> 4;
    if (e >= nelem) return;
    const int sub = (int)(t & 15), a = sub >> 1, half = sub & 1;
    double acc[4][9], fe[3];
    contract_blocks<SYM>(packet + (size_t)e * 8 * pk_len(SYM), geo + (size_t)e * 8 * 24, g0 + 24 * e, a, half, fbar,
                         acc, fe);
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) {
        const int bn = 4 * half + bb;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                Ke[576 * e + (3 * a + i) + 24 * (3 * bn + k)] = acc[bb][3 * i + k];
    }
    if (half == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) Fe[24 * e + 3 * a + i] = fe[i];
    }
}

// det(Fbar) of one Gauss point, for the ElementInversionError report.
__global__ void detF_kernel(ElementArgs a, int64_t gp, double *out) {
    const int64_t e = gp / 8;
    const int g = (int)(gp % 8);
    double Xe[24], ue[24];
    for (int n = 0; n < 8; ++n)
        for (int d = 0; d < 3; ++d) {
            const int64_t node = a.conn[8 * e + n];
            Xe[3 * n + d] = a.X[3 * node + d];
            ue[3 * n + d] = a.U[3 * node + d];
        }
    const double gq = 0.57735026918962576451;
    double dNdX[24], dj, F[9];
    ref_grads(Xe, (((g + 1) & 2) ? gq : -gq), ((g & 2) ? gq : -gq), ((g & 4) ? gq : -gq), dNdX, dj);
    defgrad(ue, dNdX, F);
    double s = 1.0;
    if (a.fbar) {
        double dNdX0[24], F0[9];
        ref_grads(Xe, 0.0, 0.0, 0.0, dNdX0, dj);
        defgrad(ue, dNdX0, F0);
        s = cbrt(det3(F0) / det3(F));
    }
    for (int k = 0; k < 9; ++k) F[k] *= s;
    *out = det3(F);
}

// gauss_stress (Model.jl:188-208): the committed Kirchhoff stress divided by det F of
// the un-barred F = I + sum_a u_a (dN_a/dX)^T at the Gauss point (1 where det F <= 0),
// one thread per Gauss point; the small-strain element stores Cauchy stress already.
__global__ __launch_bounds__(256) void cauchy_kernel(ElementArgs a, const double *__restrict__ tau,
                                                     double *__restrict__ out) {
    const int64_t gp = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gp >= 8 * a.nelem) return;
    double inv = 1.0;
    if (a.kind != 2) {
        const int64_t e = gp >> 3;
        const int g = (int)(gp & 7);
        double Xe[24], ue[24];
        for (int n = 0; n < 8; ++n) {
            const int64_t node = a.conn[8 * e + n];
            for (int d = 0; d < 3; ++d) {
                Xe[3 * n + d] = a.X[3 * node + d];
                ue[3 * n + d] = a.U[3 * node + d];
            }
        }
        const double gq = 0.57735026918962576451;
        double dNdX[24], dj, F[9];
        ref_grads(Xe, (((g + 1) & 2) ? gq : -gq), ((g & 2) ? gq : -gq), ((g & 4) ? gq : -gq), dNdX, dj);
        defgrad(ue, dNdX, F);
        const double J = det3(F);
        inv = J > 0 ? 1.0 / J : 1.0;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) out[6 * gp + k] = tau[6 * gp + k] * inv;
}

}  // namespace

void launch_cauchy_kernel(const ElementArgs &a, const double *tau, double *out, hipStream_t s) {
    const int64_t ngp = 8 * a.nelem;
    if (ngp <= 0) return;
    hipLaunchKernelGGL(cauchy_kernel, dim3((unsigned)((ngp + 255) / 256)), dim3(256), 0, s, a, tau, out);
    PF_HIP(hipGetLastError());
}

void launch_gp_kernel(const ElementArgs &a, int64_t e_begin, int64_t e_count, hipStream_t s) {
    if (e_count <= 0) return;
    const int64_t blocks = (e_count + 31) / 32;
    if (a.kind == 2)
        hipLaunchKernelGGL(gp_small_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, e_begin, e_begin + e_count);
    else if (a.kin)
        hipLaunchKernelGGL(gp_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, a, e_begin, e_begin + e_count);
    else
        hipLaunchKernelGGL(gp_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, a, e_begin, e_begin + e_count);
    PF_HIP(hipGetLastError());
}

void launch_scatter_kernel(const ElementArgs &a, const int64_t *elem_list, int64_t count,
                           const int32_t *slots, const uint64_t *first, double *Kv, double *R, hipStream_t s,
                           bool local) {
    if (count <= 0) return;
    if (packet_sym(a))
        hipLaunchKernelGGL(scatter_kernel<true>, dim3((unsigned)((count + 3) / 4)), dim3(64), 0, s, a.packet, a.geo,
                           a.g0, elem_list, count, a.conn, slots, first, Kv, R, a.fbar, (int)local);
    else
        hipLaunchKernelGGL(scatter_kernel<false>, dim3((unsigned)((count + 3) / 4)), dim3(64), 0, s, a.packet, a.geo,
                           a.g0, elem_list, count, a.conn, slots, first, Kv, R, a.fbar, (int)local);
    PF_HIP(hipGetLastError());
}

// First-touch masks of the scatter.  Elements of one colour share no node, so no two
// of them touch one K block: the block's first contribution comes from the element of
// the lowest colour that holds both its nodes.  Pass 1 takes the lowest colour per slot,
// pass 2 sets bit 8a+b of element e where (a, b) is that first touch.
__global__ void slot_fill_kernel(int64_t n, int32_t v, int32_t *__restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = v;
}
__global__ __launch_bounds__(64) void first_colour_kernel(const int64_t *__restrict__ elist,
                                                          const int32_t *__restrict__ slots, int32_t col,
                                                          int32_t *__restrict__ minc) {
    const int64_t e = elist[blockIdx.x];
    atomicMin(minc + slots[64 * e + threadIdx.x], col);
}
__global__ __launch_bounds__(64) void first_mask_kernel(const int64_t *__restrict__ elist,
                                                        const int32_t *__restrict__ slots, int32_t col,
                                                        const int32_t *__restrict__ minc,
                                                        uint64_t *__restrict__ mask) {
    const int64_t e = elist[blockIdx.x];
    const uint64_t m = __ballot(minc[slots[64 * e + threadIdx.x]] == col);
    if (threadIdx.x == 0) mask[e] = m;
}
void build_first_touch(int64_t nslots, int ncolors, const int64_t *order, const int64_t *color_ptr,
                       const int32_t *slots, uint64_t *mask, hipStream_t s) {
    dbuf<int32_t> minc(std::max<int64_t>(1, nslots));
    hipLaunchKernelGGL(slot_fill_kernel, dim3(1024), dim3(256), 0, s, nslots, INT32_MAX, minc.p);
    for (int col = 0; col < ncolors; ++col)
        if (color_ptr[col + 1] > color_ptr[col])
            hipLaunchKernelGGL(first_colour_kernel, dim3((unsigned)(color_ptr[col + 1] - color_ptr[col])), dim3(64), 0,
                               s, order + color_ptr[col], slots, col, minc.p);
    for (int col = 0; col < ncolors; ++col)
        if (color_ptr[col + 1] > color_ptr[col])
            hipLaunchKernelGGL(first_mask_kernel, dim3((unsigned)(color_ptr[col + 1] - color_ptr[col])), dim3(64), 0,
                               s, order + color_ptr[col], slots, col, minc.p, mask);
    PF_HIP(hipGetLastError());
    PF_HIP(hipStreamSynchronize(s));  // minc is freed on return
}

void launch_dense_ke_kernel(const ElementArgs &a, int64_t nelem, double *Fe, double *Ke, hipStream_t s) {
    if (nelem <= 0) return;
    const int64_t blocks = (16 * nelem + 255) / 256;
    if (packet_sym(a))
        hipLaunchKernelGGL(dense_ke_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, a.packet, a.geo, a.g0,
                           nelem, Fe, Ke, a.fbar);
    else
        hipLaunchKernelGGL(dense_ke_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, a.packet, a.geo, a.g0,
                           nelem, Fe, Ke, a.fbar);
    PF_HIP(hipGetLastError());
}

void launch_detF_kernel(const ElementArgs &a, int64_t gp, double *out, hipStream_t s) {
    hipLaunchKernelGGL(detF_kernel, dim3(1), dim3(1), 0, s, a, gp, out);
    PF_HIP(hipGetLastError());
}

}  // namespace pfem
