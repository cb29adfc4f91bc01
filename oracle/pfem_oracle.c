/*
 * pfem_oracle.c -- CPU oracle (FP64 restatement of the reference algorithm).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
 *
 * Every function cites the reference file:line (lagrangian/src/...) it restates.
 * Deliberate numerical deviation:
 *   _dhalflog_contract's off-diagonal divided difference (ln bA - ln bB)/(2(bA-bB))
 *   is evaluated as 0.5*log1p((bA-bB)/bB)/(bA-bB), which is the same quantity
 *   without the cancellation of two nearly equal logarithms; the reference's
 *   degenerate switch (EIG_TOL) and its limit 1/(bA+bB) are kept verbatim.
 * Parity pins for each part are listed in the test files under tests/.
 */
#include "pfem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EIG_TOL 1e-9 /* FiniteStrain.jl:69 */

/* ------------------------------------------------------------------ */
/* 3x3 helpers (column-major)                                           */
/* ------------------------------------------------------------------ */
#define M3(m, i, j) ((m)[(i) + 3 * (j)])

static void mm3(const double *a, const double *b, double *c) {
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += M3(a, i, k) * M3(b, k, j);
            t[i + 3 * j] = s;
        }
    memcpy(c, t, sizeof t);
}
static void tr3(const double *a, double *c) {
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[i + 3 * j] = M3(a, j, i);
    memcpy(c, t, sizeof t);
}
static double det3(const double *a) {
    return M3(a, 0, 0) * (M3(a, 1, 1) * M3(a, 2, 2) - M3(a, 1, 2) * M3(a, 2, 1)) -
           M3(a, 0, 1) * (M3(a, 1, 0) * M3(a, 2, 2) - M3(a, 1, 2) * M3(a, 2, 0)) +
           M3(a, 0, 2) * (M3(a, 1, 0) * M3(a, 2, 1) - M3(a, 1, 1) * M3(a, 2, 0));
}
static void inv3(const double *a, double *c) {
    double d = det3(a);
    double t[9];
    M3(t, 0, 0) = (M3(a, 1, 1) * M3(a, 2, 2) - M3(a, 1, 2) * M3(a, 2, 1)) / d;
    M3(t, 0, 1) = -(M3(a, 0, 1) * M3(a, 2, 2) - M3(a, 0, 2) * M3(a, 2, 1)) / d;
    M3(t, 0, 2) = (M3(a, 0, 1) * M3(a, 1, 2) - M3(a, 0, 2) * M3(a, 1, 1)) / d;
    M3(t, 1, 0) = -(M3(a, 1, 0) * M3(a, 2, 2) - M3(a, 1, 2) * M3(a, 2, 0)) / d;
    M3(t, 1, 1) = (M3(a, 0, 0) * M3(a, 2, 2) - M3(a, 0, 2) * M3(a, 2, 0)) / d;
    M3(t, 1, 2) = -(M3(a, 0, 0) * M3(a, 1, 2) - M3(a, 0, 2) * M3(a, 1, 0)) / d;
    M3(t, 2, 0) = (M3(a, 1, 0) * M3(a, 2, 1) - M3(a, 1, 1) * M3(a, 2, 0)) / d;
    M3(t, 2, 1) = -(M3(a, 0, 0) * M3(a, 2, 1) - M3(a, 0, 1) * M3(a, 2, 0)) / d;
    M3(t, 2, 2) = (M3(a, 0, 0) * M3(a, 1, 1) - M3(a, 0, 1) * M3(a, 1, 0)) / d;
    memcpy(c, t, sizeof t);
}
/* voigt_to_sym3 (FiniteStrain.jl:83-87), physical shear */
static void v2s(const double *v, double *m) {
    M3(m, 0, 0) = v[0]; M3(m, 1, 1) = v[1]; M3(m, 2, 2) = v[2];
    M3(m, 0, 1) = M3(m, 1, 0) = v[3];
    M3(m, 1, 2) = M3(m, 2, 1) = v[4];
    M3(m, 0, 2) = M3(m, 2, 0) = v[5];
}
/* sym3_to_voigt (FiniteStrain.jl:94-96) */
static void s2v(const double *m, double *v) {
    v[0] = M3(m, 0, 0); v[1] = M3(m, 1, 1); v[2] = M3(m, 2, 2);
    v[3] = M3(m, 0, 1); v[4] = M3(m, 1, 2); v[5] = M3(m, 0, 2);
}

/* ------------------------------------------------------------------ */
/* Material (Materials.jl)                                              */
/* ------------------------------------------------------------------ */
int or_material_init(or_mat *m, double E, double nu, double sy0, double Hiso,
                     double Hkin, double ssat, double delta) {
    /* guards: Materials.jl:90-103 */
    if (!(E > 0)) return 1;
    if (!(-1 < nu && nu < 0.5)) return 2;
    if (!(sy0 > 0)) return 3;
    if (!(Hiso >= 0)) return 4;
    if (!(Hkin >= 0)) return 5;
    if (!(ssat >= sy0)) return 6;
    if (!(delta >= 0)) return 7;
    if (delta > 0 && ssat > sy0 && Hkin > 0) return 8;
    m->E = E; m->nu = nu; m->sy0 = sy0; m->Hiso = Hiso; m->Hkin = Hkin;
    m->ssat = ssat; m->delta = delta;
    /* Materials.jl:104-106 */
    m->G = E / (2 * (1 + nu));
    m->K = E / (3 * (1 - 2 * nu));
    m->lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    /* elastic_matrix (Materials.jl:27-35) */
    memset(m->C, 0, sizeof m->C);
    double a = m->lam + 2 * m->G;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m->C[6 * i + j] = (i == j) ? a : m->lam;
    m->C[6 * 3 + 3] = m->C[6 * 4 + 4] = m->C[6 * 5 + 5] = m->G;
    return 0;
}

/* yield_stress / yield_slope (Materials.jl:117-127) */
static double yield_stress(const or_mat *m, double a) {
    return m->sy0 + (m->ssat - m->sy0) * (1 - exp(-m->delta * a)) + m->Hiso * a;
}
static double yield_slope(const or_mat *m, double a) {
    return m->delta * (m->ssat - m->sy0) * exp(-m->delta * a) + m->Hiso;
}

void or_return_map(const or_mat *m, const double eps[6], const double epn[6],
                   const double bn[6], double an, double sig[6], double ep[6],
                   double bt[6], double *aout, double D[36]) {
    const double G = m->G;
    double d[6], st[6];
    for (int k = 0; k < 6; ++k) d[k] = eps[k] - epn[k];
    for (int i = 0; i < 6; ++i) { /* sigma_tr = C (eps - epn)   :168 */
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s += m->C[6 * i + j] * d[j];
        st[i] = s;
    }
    double p = (st[0] + st[1] + st[2]) / 3; /* :169 */
    double xi[6];
    for (int k = 0; k < 6; ++k) xi[k] = st[k] - (k < 3 ? p : 0.0) - bn[k]; /* :170-171 */
    double nrm = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                      2 * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5])); /* devnorm :144 */
    double q = sqrt(1.5) * nrm;
    double f = q - yield_stress(m, an); /* :175 */
    if (f <= 0.0) {                     /* :177-180 elastic */
        memcpy(sig, st, 6 * sizeof(double));
        memcpy(ep, epn, 6 * sizeof(double));
        memcpy(bt, bn, 6 * sizeof(double));
        *aout = an;
        memcpy(D, m->C, 36 * sizeof(double));
        return;
    }
    const double s32 = sqrt(1.5);
    double nh[6];
    for (int k = 0; k < 6; ++k) nh[k] = xi[k] / nrm; /* :195 */
    double dg;
    if (m->delta > 0 && m->ssat > m->sy0) { /* Voce local Newton :201-213 */
        dg = f / (3 * G + m->Hkin + yield_slope(m, an));
        for (int it = 0; it < 50; ++it) {
            double Hp = yield_slope(m, an + dg);
            double r = q - (3 * G + m->Hkin) * dg - yield_stress(m, an + dg);
            double ddg = r / (3 * G + m->Hkin + Hp);
            dg += ddg;
            if (fabs(ddg) <= 1e-13 * (dg + 1e-14)) break;
        }
    } else {
        dg = f / (3 * G + m->Hiso + m->Hkin); /* :218 */
    }
    for (int k = 0; k < 6; ++k) sig[k] = st[k] - (2 * G * dg * s32) * nh[k]; /* :222 */
    for (int k = 0; k < 6; ++k)                                              /* :226-227 */
        ep[k] = epn[k] + (dg * s32) * (k < 3 ? nh[k] : 2 * nh[k]);
    for (int k = 0; k < 6; ++k) bt[k] = bn[k] + (2.0 / 3.0) * m->Hkin * (dg * s32) * nh[k]; /* :230 */
    double anew = an + dg;
    *aout = anew;
    /* consistent tangent :244-247 */
    double Hp = yield_slope(m, anew);
    double b0 = 2 * G * dg / q;
    double g0 = 2 * G / (3 * G + m->Hkin + Hp);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double isym = (i == j) ? (i < 3 ? 1.0 : 0.5) : 0.0;
            double oxo = (i < 3 && j < 3) ? 1.0 : 0.0;
            double idev = isym - oxo / 3;
            D[6 * i + j] = m->C[6 * i + j] - (3 * G * b0) * idev + (3 * G * (b0 - g0)) * (nh[i] * nh[j]);
        }
}

/* ------------------------------------------------------------------ */
/* Symmetric 3x3 eigen-decomposition: cyclic Jacobi rotations           */
/* (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.5.2/8.5.3).  */
/* Eigen order/sign differs from StaticArrays' analytic solver; every   */
/* quantity built from it is an isotropic tensor function, invariant to */
/* that choice.                                                         */
/* ------------------------------------------------------------------ */
void or_eig_sym3(const double Ain[9], double w[3], double V[9]) {
    double a[9];
    memcpy(a, Ain, sizeof a);
    for (int k = 0; k < 9; ++k) V[k] = (k % 4 == 0) ? 1.0 : 0.0;
    static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = M3(a, 0, 1) * M3(a, 0, 1) + M3(a, 0, 2) * M3(a, 0, 2) + M3(a, 1, 2) * M3(a, 1, 2);
        double dia = M3(a, 0, 0) * M3(a, 0, 0) + M3(a, 1, 1) * M3(a, 1, 1) + M3(a, 2, 2) * M3(a, 2, 2);
        if (off <= 1e-36 * dia || off == 0.0) break;
        for (int r = 0; r < 3; ++r) {
            int p = PQ[r][0], q = PQ[r][1];
            double apq = M3(a, p, q);
            if (apq == 0.0) continue;
            double theta = (M3(a, q, q) - M3(a, p, p)) / (2 * apq);
            double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
            double c = 1 / sqrt(t * t + 1), s = t * c;
            /* A <- J^T A J with J the (p,q) rotation */
            for (int k = 0; k < 3; ++k) {
                double akp = M3(a, k, p), akq = M3(a, k, q);
                M3(a, k, p) = c * akp - s * akq;
                M3(a, k, q) = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {
                double apk = M3(a, p, k), aqk = M3(a, q, k);
                M3(a, p, k) = c * apk - s * aqk;
                M3(a, q, k) = s * apk + c * aqk;
            }
            M3(a, p, q) = M3(a, q, p) = 0.0;
            for (int k = 0; k < 3; ++k) {
                double vkp = M3(V, k, p), vkq = M3(V, k, q);
                M3(V, k, p) = c * vkp - s * vkq;
                M3(V, k, q) = s * vkp + c * vkq;
            }
        }
    }
    w[0] = M3(a, 0, 0); w[1] = M3(a, 1, 1); w[2] = M3(a, 2, 2);
    /* ascending order (insertion sort with vector columns) */
    for (int i = 1; i < 3; ++i)
        for (int j = i; j > 0 && w[j - 1] > w[j]; --j) {
            double tw = w[j]; w[j] = w[j - 1]; w[j - 1] = tw;
            for (int k = 0; k < 3; ++k) {
                double tv = M3(V, k, j); M3(V, k, j) = M3(V, k, j - 1); M3(V, k, j - 1) = tv;
            }
        }
}

/* ------------------------------------------------------------------ */
/* Finite-strain wrap (FiniteStrain.jl)                                 */
/* ------------------------------------------------------------------ */
void or_finite_kinematics(const double F[9], const double cpi[6], or_kin *k) {
    /* FiniteStrain.jl:148-170 */
    double J = det3(F);
    k->J = J;
    if (J <= 0) {
        memset(k->ee, 0, sizeof k->ee);
        k->b[0] = k->b[1] = k->b[2] = 1.0;
        for (int i = 0; i < 9; ++i) k->n[i] = k->Finv[i] = (i % 4 == 0) ? 1.0 : 0.0;
        k->ok = 0;
        return;
    }
    double C[9], Ft[9], be[9];
    v2s(cpi, C);
    tr3(F, Ft);
    mm3(F, C, be);
    mm3(be, Ft, be);
    /* symmetrize exactly (Symmetric() reads the upper triangle) */
    M3(be, 1, 0) = M3(be, 0, 1); M3(be, 2, 0) = M3(be, 0, 2); M3(be, 2, 1) = M3(be, 1, 2);
    or_eig_sym3(be, k->b, k->n);
    double ll[3];
    for (int a = 0; a < 3; ++a) ll[a] = 0.5 * log(k->b[a]);
    double em[9] = {0};
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M3(em, i, j) += ll[a] * M3(k->n, i, a) * M3(k->n, j, a);
    k->ee[0] = M3(em, 0, 0); k->ee[1] = M3(em, 1, 1); k->ee[2] = M3(em, 2, 2);
    k->ee[3] = 2 * M3(em, 0, 1); k->ee[4] = 2 * M3(em, 1, 2); k->ee[5] = 2 * M3(em, 0, 2);
    inv3(F, k->Finv);
    k->ok = 1;
}

/* polar_RU (FiniteStrain.jl:190-205): F = R U, U = (F^T F)^(1/2), R = F U^-1. */
void or_polar_RU(const double F[9], double R[9], double U[9]) {
    double Ft[9], C[9], w[3], V[9], sq[3], isq[3], Ui[9] = {0};
    tr3(F, Ft);
    mm3(Ft, F, C);
    M3(C, 1, 0) = M3(C, 0, 1); M3(C, 2, 0) = M3(C, 0, 2); M3(C, 2, 1) = M3(C, 1, 2);
    or_eig_sym3(C, w, V);
    for (int a = 0; a < 3; ++a) {
        sq[a] = sqrt(w[a] > 0.0 ? w[a] : 0.0); /* floor at 0 (:195-197) */
        isq[a] = sq[a] > 0 ? 1.0 / sq[a] : 0.0;
    }
    memset(U, 0, 9 * sizeof(double));
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                M3(U, i, j) += sq[a] * M3(V, i, a) * M3(V, j, a);
                M3(Ui, i, j) += isq[a] * M3(V, i, a) * M3(V, j, a);
            }
    mm3(F, Ui, R);
}

/* dR_polar (FiniteStrain.jl:215-228): Omega U + U Omega = R^T dF - dF^T R solved in
 * U's eigenbasis, dR = R Omega. */
void or_dR_polar(const double R[9], const double U[9], const double dF[9], double dR[9]) {
    double Rt[9], A[9], At[9], rhs[9], w[3], V[9], Vt[9], rb[9], Ob[9], t[9];
    tr3(R, Rt);
    mm3(Rt, dF, A);
    tr3(A, At);
    for (int i = 0; i < 9; ++i) rhs[i] = A[i] - At[i];
    double Us[9];
    memcpy(Us, U, sizeof Us);
    M3(Us, 1, 0) = M3(Us, 0, 1); M3(Us, 2, 0) = M3(Us, 0, 2); M3(Us, 2, 1) = M3(Us, 1, 2);
    or_eig_sym3(Us, w, V);
    tr3(V, Vt);
    mm3(Vt, rhs, t);
    mm3(t, V, rb);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M3(Ob, i, j) = M3(rb, i, j) / (w[i] + w[j]);
    mm3(V, Ob, t);
    mm3(t, Vt, t);
    mm3(R, t, dR);
}

/* dtau_dbeta (FiniteStrain.jl:245-262), row-major 6x6; zero in the elastic branch. */
void or_dtau_dbeta(const or_mat *m, const double ee[6], const double bsp[6], double an,
                   double out[36]) {
    const double G = m->G, H = m->Hiso + m->Hkin, s32 = sqrt(1.5);
    double st[6], xi[6];
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s += m->C[6 * i + j] * ee[j];
        st[i] = s;
    }
    double p = (st[0] + st[1] + st[2]) / 3;
    for (int k = 0; k < 6; ++k) xi[k] = st[k] - (k < 3 ? p : 0.0) - bsp[k];
    double nrm = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                      2 * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
    double Y = m->sy0 + m->Hiso * an;
    double f = s32 * nrm - Y;
    memset(out, 0, 36 * sizeof(double));
    if (f <= 0.0) return;
    double dg = f / (3 * G + H);
    double a = 2 * G * s32 * dg / nrm;
    double da = 2 * G * s32 / (3 * G + H) * (Y / (nrm * nrm));
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
            out[6 * i + j] = (i == j ? a : 0.0) + xi[i] * (da / nrm) * ((j < 3 ? 1.0 : 2.0) * xi[j]);
}

void or_finite_stress_update(const or_mat *m, const or_kin *k, const double F[9], const double epn[6],
                             const double bref_n[6], double an, double tau[6], double ep[6],
                             double bref_new[6], double *aout, double D[36], double cpi_new[6],
                             double bsp_out[6], double R_out[9], double U_out[9]) {
    /* FiniteStrain.jl:292-346 */
    double Z6[6] = {0}, dep[6], bsp[6] = {0}, bsp_new[6], R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1},
           U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const int kin = m->Hkin > 0;
    if (kin) { /* :312-314: beta_sp = R beta_ref R^T */
        double B[9], t[9], Rt[9];
        or_polar_RU(F, R, U);
        v2s(bref_n ? bref_n : Z6, B);
        tr3(R, Rt);
        mm3(R, B, t);
        mm3(t, Rt, t);
        s2v(t, bsp);
    }
    or_return_map(m, k->ee, Z6, bsp, an, tau, dep, bsp_new, aout, D); /* :318 */
    if (bref_new) {                                                  /* :320 */
        if (kin) {
            double B[9], t[9], Rt[9];
            v2s(bsp_new, B);
            tr3(R, Rt);
            mm3(Rt, B, t);
            mm3(t, R, t);
            s2v(t, bref_new);
        } else {
            memcpy(bref_new, bsp_new, 6 * sizeof(double));
        }
    }
    if (bsp_out) memcpy(bsp_out, bsp, sizeof bsp);
    if (R_out) memcpy(R_out, R, sizeof R);
    if (U_out) memcpy(U_out, U, sizeof U);
    for (int i = 0; i < 6; ++i) ep[i] = epn[i] + dep[i];        /* :329 */
    if (!cpi_new) return;
    double ev[6], et[9], w[3], V[9], be[9] = {0}, FiT[9], t[9];
    for (int i = 0; i < 6; ++i) ev[i] = k->ee[i] - dep[i];      /* :334 */
    /* 2 * eps_e tensor (:335-338) */
    M3(et, 0, 0) = 2 * ev[0]; M3(et, 1, 1) = 2 * ev[1]; M3(et, 2, 2) = 2 * ev[2];
    M3(et, 0, 1) = M3(et, 1, 0) = 2 * (0.5 * ev[3]);
    M3(et, 1, 2) = M3(et, 2, 1) = 2 * (0.5 * ev[4]);
    M3(et, 0, 2) = M3(et, 2, 0) = 2 * (0.5 * ev[5]);
    or_eig_sym3(et, w, V);
    for (int a = 0; a < 3; ++a) { /* :339-341 */
        double ex = exp(w[a]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M3(be, i, j) += ex * M3(V, i, a) * M3(V, j, a);
    }
    tr3(k->Finv, FiT); /* :342 */
    mm3(k->Finv, be, t);
    mm3(t, FiT, t);
    s2v(t, cpi_new);
}

/* _dhalflog_contract (FiniteStrain.jl:463-482) */
static void dhalflog_contract(const double *Q, const double *b, const double *H, double *out) {
    double Qt[9], Hp[9], Yp[9], t[9];
    tr3(Q, Qt);
    mm3(Qt, H, t);
    mm3(t, Q, Hp);
    for (int A = 0; A < 3; ++A)
        for (int B = 0; B < 3; ++B) {
            double g;
            if (A == B)
                g = 1.0 / (2 * b[A]);
            else if (fabs(b[A] - b[B]) < EIG_TOL * (fabs(b[A]) + fabs(b[B]) + 1.0))
                g = 1.0 / (b[A] + b[B]);
            else {
                double d = b[A] - b[B]; /* = (0.5 ln bA - 0.5 ln bB) / (bA - bB) */
                g = 0.5 * log1p(d / b[B]) / d;
            }
            M3(Yp, A, B) = g * M3(Hp, A, B);
        }
    mm3(Q, Yp, t);
    mm3(t, Qt, out);
}

void or_dPdF(const or_kin *k, const double cpi[6], const double D[36],
             const double tau[6], const double F[9], const double bref[6], const double R[9],
             const double U[9], const double dtdb[36], double A[81]) {
    /* FiniteStrain.jl:503-547; the back-stress rotation term (:533-537) is active when
     * bref, R, U and dtdb are all given */
    double C[9], Ft[9], FiT[9], T[9], B[9], Rt[9];
    const int kin = bref && R && U && dtdb;
    if (kin) {
        v2s(bref, B);
        tr3(R, Rt);
    }
    v2s(cpi, C);
    tr3(F, Ft);
    tr3(k->Finv, FiT);
    v2s(tau, T);
    for (int col = 0; col < 9; ++col) {
        int p = col % 3, q = col / 3;
        double dF[9] = {0}, dFt[9], a1[9], a2[9], dbe[9], de[9];
        M3(dF, p, q) = 1.0;
        tr3(dF, dFt);
        mm3(dF, C, a1); mm3(a1, Ft, a1); /* dF Cpi F^T */
        mm3(F, C, a2); mm3(a2, dFt, a2); /* F Cpi dF^T */
        for (int i = 0; i < 9; ++i) dbe[i] = a1[i] + a2[i];
        dhalflog_contract(k->n, k->b, dbe, de);
        double ev[6] = {M3(de, 0, 0), M3(de, 1, 1), M3(de, 2, 2),
                        2 * M3(de, 0, 1), 2 * M3(de, 1, 2), 2 * M3(de, 0, 2)}; /* _eng_voigt :485 */
        double dtv[6], dt[9];
        for (int i = 0; i < 6; ++i) {
            double s = 0.0;
            for (int j = 0; j < 6; ++j) s += D[6 * i + j] * ev[j];
            dtv[i] = s;
        }
        if (kin) { /* d beta_sp = dR B R^T + R B dR^T; d tau += dtdb d beta_sp */
            double dR[9], t1[9], t2[9], dRt[9], dbs[9], dbv[6];
            or_dR_polar(R, U, dF, dR);
            mm3(dR, B, t1); mm3(t1, Rt, t1);
            tr3(dR, dRt);
            mm3(R, B, t2); mm3(t2, dRt, t2);
            for (int i = 0; i < 9; ++i) dbs[i] = t1[i] + t2[i];
            s2v(dbs, dbv);
            for (int i = 0; i < 6; ++i) {
                double s = 0.0;
                for (int j = 0; j < 6; ++j) s += dtdb[6 * i + j] * dbv[j];
                dtv[i] += s;
            }
        }
        v2s(dtv, dt);
        double b1[9], b2[9], dFiT[9], dP[9];
        mm3(FiT, dFt, b1); mm3(b1, FiT, b1);
        for (int i = 0; i < 9; ++i) dFiT[i] = -b1[i]; /* :539 */
        mm3(dt, FiT, b1);
        mm3(T, dFiT, b2);
        for (int i = 0; i < 9; ++i) dP[i] = b1[i] + b2[i]; /* :540 */
        for (int i = 0; i < 9; ++i) A[i + 9 * col] = dP[i];
    }
}

/* ------------------------------------------------------------------ */
/* Hex8 element (Elements.jl)                                           */
/* ------------------------------------------------------------------ */
static const double NODE_NAT[8][3] = {/* Elements.jl:26-30 */
                                      {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                      {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
static const double GSGN[8][3] = {/* Elements.jl:34-42 (times 1/sqrt(3)) */
                                  {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void or_hex8_dshape(const double xi[3], double dN[24]) { /* Elements.jl:62-75 */
    for (int a = 0; a < 8; ++a) {
        double xa = NODE_NAT[a][0], ya = NODE_NAT[a][1], za = NODE_NAT[a][2];
        dN[3 * a + 0] = 0.125 * xa * (1 + ya * xi[1]) * (1 + za * xi[2]);
        dN[3 * a + 1] = 0.125 * ya * (1 + xa * xi[0]) * (1 + za * xi[2]);
        dN[3 * a + 2] = 0.125 * za * (1 + xa * xi[0]) * (1 + ya * xi[1]);
    }
}

void or_ref_grads(const double Xe[24], const double xi[3], double dNdX[24], double *detJ) {
    /* jacobian J = Xe^T dN (Elements.jl:89); dNdX = dN J^{-1} (:159) */
    double dN[24], J[9] = {0}, Ji[9];
    or_hex8_dshape(xi, dN);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s += Xe[3 * a + i] * dN[3 * a + j];
            M3(J, i, j) = s;
        }
    *detJ = det3(J);
    inv3(J, Ji);
    for (int a = 0; a < 8; ++a)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += dN[3 * a + k] * M3(Ji, k, j);
            dNdX[3 * a + j] = s;
        }
}

/* deformation_gradient (FiniteStrain.jl:107-119) */
static void defgrad(const double *ue, const double *dNdX, double *F) {
    double H[9] = {0};
    for (int a = 0; a < 8; ++a)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) M3(H, i, j) += ue[3 * a + i] * dNdX[3 * a + j];
    for (int i = 0; i < 9; ++i) F[i] = H[i] + ((i % 4 == 0) ? 1.0 : 0.0);
}

/* G (9x24) applied implicitly: G[(j*3+i), (3a+k)] = delta_ik dNdX[a][j]   (Elements.jl:500-510) */

int or_element_finite(int fbar, const or_mat *m, const double Xe[24],
                      const double ue[24], double *epsp, double *beta, double *abar,
                      double *cpinv, double *sig, int commit, double Fe[24],
                      double Ke[576], int *fail_g, double *fail_J) {
    const double gq = 1.0 / sqrt(3.0);
    double dNdX0[24], detJ0, F0[9], J0 = 1.0, F0i[9], g0[24] = {0};
    if (fbar) { /* Elements.jl:421-424, centroid quantities */
        const double z[3] = {0, 0, 0};
        or_ref_grads(Xe, z, dNdX0, &detJ0);
        defgrad(ue, dNdX0, F0);
        J0 = det3(F0);
        inv3(F0, F0i);
        /* g0 = G0^T vec(F0^{-T}) (Elements.jl:528): g0[a,k] = sum_j dNdX0[a,j] F0^{-T}[k,j] */
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int j = 0; j < 3; ++j) s += dNdX0[3 * a + j] * M3(F0i, j, k);
                g0[3 * a + k] = s;
            }
    }
    memset(Fe, 0, 24 * sizeof(double));
    memset(Ke, 0, 576 * sizeof(double));
    for (int g = 0; g < 8; ++g) { /* Elements.jl:428-492 */
        double xi[3] = {GSGN[g][0] * gq, GSGN[g][1] * gq, GSGN[g][2] * gq};
        double dNdX[24], detJ, F[9], Fb[9];
        or_ref_grads(Xe, xi, dNdX, &detJ);
        double w = detJ * 1.0;
        defgrad(ue, dNdX, F);
        double Jg = det3(F);
        double scale = fbar ? cbrt(J0 / Jg) : 1.0; /* :437 */
        for (int i = 0; i < 9; ++i) Fb[i] = fbar ? scale * F[i] : F[i];
        double *cp = cpinv + 6 * g;
        or_kin kin;
        or_finite_kinematics(Fb, cp, &kin);
        if (!kin.ok) {
            if (fail_g) *fail_g = g;
            if (fail_J) *fail_J = kin.J;
            return 1; /* :447 */
        }
        double tau[6], epn[6], ep[6], an = abar[g], anew, D[36], cn[6], bn[6], bnew[6], bsp[6],
            Rp[9], Up[9], dtdb[36];
        memcpy(epn, epsp + 6 * g, sizeof epn);
        memcpy(bn, beta + 6 * g, sizeof bn);
        or_finite_stress_update(m, &kin, Fb, epn, bn, an, tau, ep, bnew, &anew, D, commit ? cn : NULL,
                                bsp, Rp, Up);
        double T[9], FiT[9], P[9], A[81];
        v2s(tau, T);
        tr3(kin.Finv, FiT);
        mm3(T, FiT, P); /* first_piola :560 ; P column-major == _p9 */
        if (m->Hkin > 0) { /* Elements.jl:467-469 */
            or_dtau_dbeta(m, kin.ee, bsp, an, dtdb);
            or_dPdF(&kin, cp, D, tau, Fb, bn, Rp, Up, dtdb, A);
        } else {
            or_dPdF(&kin, cp, D, tau, Fb, NULL, NULL, NULL, NULL, A);
        }
        /* Fe += G^T P w   (:470) */
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i) {
                double s = 0.0;
                for (int j = 0; j < 3; ++j) s += dNdX[3 * a + j] * P[i + 3 * j];
                Fe[3 * a + i] += s * w;
            }
        /* Geff (9x24) = scale*G + (scale/3) vec(F) (g0 - gJ)^T   (:524-531) */
        double Geff[9 * 24];
        for (int c = 0; c < 24; ++c) {
            int a = c / 3, k = c % 3;
            for (int r = 0; r < 9; ++r) {
                int i = r % 3, j = r / 3;
                Geff[r + 9 * c] = (i == k) ? dNdX[3 * a + j] : 0.0;
            }
        }
        if (fbar) {
            double Fi[9], gJ[24];
            inv3(F, Fi);
            for (int a = 0; a < 8; ++a)
                for (int k = 0; k < 3; ++k) {
                    double s = 0.0;
                    for (int j = 0; j < 3; ++j) s += dNdX[3 * a + j] * M3(Fi, j, k);
                    gJ[3 * a + k] = s;
                }
            for (int c = 0; c < 24; ++c)
                for (int r = 0; r < 9; ++r)
                    Geff[r + 9 * c] = scale * Geff[r + 9 * c] + (scale / 3) * (F[r] * (g0[c] - gJ[c]));
        }
        /* AG = A * Geff (9x24);  Ke += G^T AG w  (:475/:477) */
        double AG[9 * 24];
        for (int c = 0; c < 24; ++c)
            for (int r = 0; r < 9; ++r) {
                double s = 0.0;
                for (int l = 0; l < 9; ++l) s += A[r + 9 * l] * Geff[l + 9 * c];
                AG[r + 9 * c] = s;
            }
        for (int c = 0; c < 24; ++c)
            for (int a = 0; a < 8; ++a)
                for (int i = 0; i < 3; ++i) {
                    double s = 0.0;
                    for (int j = 0; j < 3; ++j) s += dNdX[3 * a + j] * AG[(i + 3 * j) + 9 * c];
                    Ke[(3 * a + i) + 24 * c] += s * w;
                }
        memcpy(sig + 6 * g, tau, 6 * sizeof(double)); /* :481-482 */
        if (commit) {                                  /* :483-491 */
            memcpy(epsp + 6 * g, ep, 6 * sizeof(double));
            memcpy(beta + 6 * g, bnew, 6 * sizeof(double));
            abar[g] = anew;
            memcpy(cp, cn, 6 * sizeof(double));
        }
    }
    return 0;
}

/* Small-strain Hex8 (element_force_tangent! Elements.jl:323-360 with the B matrices of
 * element_geometry :130-144 and bmatrix :97-118): eps = B ue (engineering shear),
 * return_map with the committed eps_p, beta, abar; Fe += B^T sigma w, Ke += B^T D B w. */
static const int VOIGT_OF[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
void or_element_small(const or_mat *m, const double Xe[24], const double ue[24], double *epsp,
                      double *beta, double *abar, double *sig, int commit, double Fe[24],
                      double Ke[576]) {
    const double gq = 1.0 / sqrt(3.0);
    memset(Fe, 0, 24 * sizeof(double));
    memset(Ke, 0, 576 * sizeof(double));
    for (int g = 0; g < 8; ++g) {
        double xi[3] = {GSGN[g][0] * gq, GSGN[g][1] * gq, GSGN[g][2] * gq};
        double dNdX[24], detJ, H[9] = {0};
        or_ref_grads(Xe, xi, dNdX, &detJ);
        const double w = detJ * 1.0;
        for (int a = 0; a < 8; ++a)
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) M3(H, i, j) += ue[3 * a + i] * dNdX[3 * a + j];
        double eps[6] = {M3(H, 0, 0), M3(H, 1, 1), M3(H, 2, 2), M3(H, 0, 1) + M3(H, 1, 0),
                         M3(H, 1, 2) + M3(H, 2, 1), M3(H, 0, 2) + M3(H, 2, 0)};
        double s[6], ep[6], bt[6], an, D[36];
        or_return_map(m, eps, epsp + 6 * g, beta + 6 * g, abar[g], s, ep, bt, &an, D);
        /* B[v][(a,i)] = dNdX[a][j] for the one j with VOIGT_OF[i][j] == v */
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i) {
                double f = 0.0;
                for (int j = 0; j < 3; ++j) f += dNdX[3 * a + j] * s[VOIGT_OF[i][j]];
                Fe[3 * a + i] += f * w;
                for (int b = 0; b < 8; ++b)
                    for (int k = 0; k < 3; ++k) {
                        double kv = 0.0;
                        for (int j = 0; j < 3; ++j)
                            for (int l = 0; l < 3; ++l)
                                kv += dNdX[3 * a + j] * D[6 * VOIGT_OF[i][j] + VOIGT_OF[k][l]] * dNdX[3 * b + l];
                        Ke[(3 * a + i) + 24 * (3 * b + k)] += kv * w;
                    }
            }
        memcpy(sig + 6 * g, s, 6 * sizeof(double));
        if (commit) {
            memcpy(epsp + 6 * g, ep, 6 * sizeof(double));
            memcpy(beta + 6 * g, bt, 6 * sizeof(double));
            abar[g] = an;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Mesh (Mesh.jl:30-59) + seeded jitter                                 */
/* ------------------------------------------------------------------ */
uint64_t or_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double or_jitter_u01(uint64_t seed, int64_t node, int comp) {
    uint64_t h = or_splitmix64(or_splitmix64(seed) ^ (uint64_t)(3 * node + comp));
    return (double)(h >> 11) * 0x1.0p-53;
}

void or_box_mesh(double lx, double ly, double lz, int nx, int ny, int nz,
                 double jit, uint64_t seed, double *nodes, int64_t *conn) {
    int64_t npx = nx + 1, npy = ny + 1;
    for (int64_t k = 0; k <= nz; ++k)
        for (int64_t j = 0; j <= ny; ++j)
            for (int64_t i = 0; i <= nx; ++i) {
                int64_t n = (k * npy + j) * npx + i;
                double x = lx * (double)i / nx, y = ly * (double)j / ny, z = lz * (double)k / nz;
                int interior = (i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz);
                if (jit != 0.0 && interior) {
                    x += (2 * or_jitter_u01(seed, n, 0) - 1) * jit * (lx / nx);
                    y += (2 * or_jitter_u01(seed, n, 1) - 1) * jit * (ly / ny);
                    z += (2 * or_jitter_u01(seed, n, 2) - 1) * jit * (lz / nz);
                }
                nodes[3 * n + 0] = x; nodes[3 * n + 1] = y; nodes[3 * n + 2] = z;
            }
    int64_t e = 0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i, ++e) {
#define NID(ii, jj, kk) (((kk) * npy + (jj)) * npx + (ii))
                int64_t *c = conn + 8 * e;
                c[0] = NID(i, j, k);         c[1] = NID(i + 1, j, k);
                c[2] = NID(i + 1, j + 1, k); c[3] = NID(i, j + 1, k);
                c[4] = NID(i, j, k + 1);     c[5] = NID(i + 1, j, k + 1);
                c[6] = NID(i + 1, j + 1, k + 1); c[7] = NID(i, j + 1, k + 1);
#undef NID
            }
}

/* ------------------------------------------------------------------ */
/* Sparsity (Assembly.jl:62-162) and colours                            */
/* ------------------------------------------------------------------ */
static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* node adjacency as CSR: sorted, unique (Assembly.jl:62-82) */
static int node_adjacency(int64_t nnodes, int64_t nelem, const int64_t *conn,
                          int64_t **aptr_out, int64_t **aidx_out) {
    int64_t *start = calloc(nnodes + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nelem; ++e)
        for (int a = 0; a < 8; ++a) start[conn[8 * e + a] + 1] += 8;
    for (int64_t n = 0; n < nnodes; ++n) start[n + 1] += start[n];
    int64_t *tmp = malloc(sizeof(int64_t) * (start[nnodes] ? start[nnodes] : 1));
    int64_t *fill = malloc(sizeof(int64_t) * (nnodes + 1));
    memcpy(fill, start, sizeof(int64_t) * (nnodes + 1));
    for (int64_t e = 0; e < nelem; ++e)
        for (int a = 0; a < 8; ++a) {
            int64_t na = conn[8 * e + a];
            for (int b = 0; b < 8; ++b) tmp[fill[na]++] = conn[8 * e + b];
        }
    int64_t *aptr = malloc(sizeof(int64_t) * (nnodes + 1));
    aptr[0] = 0;
    for (int64_t n = 0; n < nnodes; ++n) { /* sort! + unique! per node */
        int64_t s = start[n], len = start[n + 1] - start[n];
        qsort(tmp + s, len, sizeof(int64_t), cmp_i64);
        int64_t u = 0;
        for (int64_t t = 0; t < len; ++t)
            if (u == 0 || tmp[s + t] != tmp[s + u - 1]) tmp[s + u++] = tmp[s + t];
        aptr[n + 1] = aptr[n] + u;
    }
    int64_t *aidx = malloc(sizeof(int64_t) * (aptr[nnodes] ? aptr[nnodes] : 1));
    for (int64_t n = 0; n < nnodes; ++n)
        memcpy(aidx + aptr[n], tmp + start[n], sizeof(int64_t) * (aptr[n + 1] - aptr[n]));
    free(start); free(tmp); free(fill);
    *aptr_out = aptr;
    *aidx_out = aidx;
    return 0;
}

int or_build_csc(int64_t nnodes, int64_t nelem, const int64_t *conn,
                 int64_t *colptr, int64_t *rowval, int64_t *nnz_out) {
    int64_t *aptr, *aidx;
    node_adjacency(nnodes, nelem, conn, &aptr, &aidx);
    /* count (Assembly.jl:107-128) */
    colptr[0] = 0;
    for (int64_t jn = 0; jn < nnodes; ++jn) {
        int64_t ncol = 3 * (aptr[jn + 1] - aptr[jn]);
        for (int c = 0; c < 3; ++c) colptr[3 * jn + c + 1] = colptr[3 * jn + c] + ncol;
    }
    *nnz_out = colptr[3 * nnodes];
    if (rowval) { /* fill (Assembly.jl:132-146) */
        for (int64_t jn = 0; jn < nnodes; ++jn)
            for (int c = 0; c < 3; ++c) {
                int64_t p = colptr[3 * jn + c];
                for (int64_t t = aptr[jn]; t < aptr[jn + 1]; ++t) {
                    int64_t r0 = 3 * aidx[t];
                    rowval[p++] = r0; rowval[p++] = r0 + 1; rowval[p++] = r0 + 2;
                }
            }
    }
    free(aptr); free(aidx);
    return 0;
}

int or_greedy_colors(int64_t nnodes, int64_t nelem, const int64_t *conn, int *color) {
    uint64_t *mask = calloc(nnodes, sizeof(uint64_t));
    int ncol = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        uint64_t forb = 0;
        for (int a = 0; a < 8; ++a) forb |= mask[conn[8 * e + a]];
        int c = 0;
        while (c < 64 && ((forb >> c) & 1)) ++c;
        if (c == 64) { free(mask); return -1; }
        color[e] = c;
        if (c + 1 > ncol) ncol = c + 1;
        for (int a = 0; a < 8; ++a) mask[conn[8 * e + a]] |= (1ull << c);
    }
    free(mask);
    return ncol;
}

/* _find_nz binary search (Assembly.jl:210-224) */
static int64_t find_nz(const int64_t *rowval, int64_t lo, int64_t hi, int64_t grow) {
    while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        int64_t rv = rowval[mid];
        if (rv == grow) return mid;
        if (rv < grow) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

int or_assemble(int kind, const or_mat *mats, const int32_t *matid, int64_t nnodes, int64_t nelem,
                const double *nodes, const int64_t *conn, const int64_t *colptr,
                const int64_t *rowval, int ncolors, const int64_t *colour_ptr,
                const int64_t *colour_order, const double *U, double *epsp, double *beta,
                double *abar, double *cpinv, double *sig, int commit,
                double *nzval, double *R, int nthreads, int64_t *fail_e,
                int *fail_g, double *fail_J) {
    (void)nelem;
    int64_t ndof = 3 * nnodes;
    memset(nzval, 0, sizeof(double) * colptr[ndof]); /* Assembly.jl:361-362 */
    memset(R, 0, sizeof(double) * ndof);
    int failed = 0;
    int64_t fe = -1; int fg = -1; double fJ = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    for (int c = 0; c < ncolors && !failed; ++c) { /* Assembly.jl:378-383 */
        int64_t b = colour_ptr[c], eend = colour_ptr[c + 1];
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t t = b; t < eend; ++t) {
            int64_t e = colour_order[t];
            const int64_t *ce = conn + 8 * e;
            double Xe[24], ue[24], Fe[24], Ke[576];
            int64_t ed[24];
            for (int a = 0; a < 8; ++a)
                for (int d = 0; d < 3; ++d) {
                    Xe[3 * a + d] = nodes[3 * ce[a] + d];
                    ed[3 * a + d] = 3 * ce[a] + d;
                    ue[3 * a + d] = U[ed[3 * a + d]];
                }
            int g = 0; double J = 0.0;
            const or_mat *m = matid ? mats + matid[e] : mats; /* per-element phase (two-phase blocks) */
            int st = 0;
            if (kind == 2) /* Hex8Small (Assembly.jl:287-292) */
                or_element_small(m, Xe, ue, epsp + 48 * e, beta + 48 * e, abar + 8 * e, sig + 48 * e,
                                 commit, Fe, Ke);
            else
                st = or_element_finite(kind == 1, m, Xe, ue, epsp + 48 * e, beta + 48 * e, abar + 8 * e,
                                       cpinv + 48 * e, sig + 48 * e, commit, Fe, Ke, &g, &J);
            if (st) {
#pragma omp critical
                {
                    if (!failed || e < fe) { failed = 1; fe = e; fg = g; fJ = J; }
                }
                continue;
            }
            for (int cc = 0; cc < 24; ++cc) { /* scatter (Assembly.jl:401-413) */
                int64_t gcol = ed[cc];
                int64_t cs = colptr[gcol], ce2 = colptr[gcol + 1] - 1;
                for (int r = 0; r < 24; ++r) {
                    int64_t idx = find_nz(rowval, cs, ce2, ed[r]);
                    nzval[idx] += Ke[r + 24 * cc];
                }
            }
            for (int i = 0; i < 24; ++i) R[ed[i]] += Fe[i];
        }
    }
    if (failed) {
        if (fail_e) *fail_e = fe;
        if (fail_g) *fail_g = fg;
        if (fail_J) *fail_J = fJ;
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* impose_dirichlet! (BoundaryConditions.jl:62-109)                     */
/* ------------------------------------------------------------------ */
void or_impose_dirichlet(int64_t n, const int64_t *colptr, const int64_t *rowval,
                         double *nzval, double *R, int64_t nbc, const int64_t *dofs,
                         const double *values, int ramp, double lambda,
                         const double *U) {
    char *isbc = calloc(n, 1);
    double *dg = calloc(n, sizeof(double));
    for (int64_t i = 0; i < nbc; ++i) {
        int64_t d = dofs[i];
        double g = ramp ? lambda * values[i] : values[i];
        isbc[d] = 1;
        dg[d] = g - U[d];
    }
    for (int64_t col = 0; col < n; ++col)
        for (int64_t k = colptr[col]; k < colptr[col + 1]; ++k) {
            int64_t row = rowval[k];
            if (isbc[col] && !isbc[row]) R[row] += nzval[k] * dg[col];
        }
    for (int64_t col = 0; col < n; ++col)
        for (int64_t k = colptr[col]; k < colptr[col + 1]; ++k) {
            int64_t row = rowval[k];
            if (isbc[col] || isbc[row]) nzval[k] = (row == col && isbc[row]) ? 1.0 : 0.0;
        }
    for (int64_t i = 0; i < nbc; ++i) R[dofs[i]] = -dg[dofs[i]];
    free(isbc);
    free(dg);
}

/* ------------------------------------------------------------------ */
/* gauss_stress (Model.jl:188-208): committed Kirchhoff stress tau -> */
/* Cauchy sigma = tau / J with J = det F of the un-barred              */
/* F = I + grad U (_defgrad, Model.jl:212-222) at each Gauss point,    */
/* 1/J replaced by 1 where J <= 0.  kind 2 (small strain) copies tau.   */
/* ------------------------------------------------------------------ */
void or_gauss_stress(int kind, int64_t nelem, const double *nodes, const int64_t *conn,
                     const double *U, const double *tau, double *sig) {
    for (int64_t e = 0; e < nelem; ++e) {
        double Xe[24], ue[24];
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) {
                Xe[3 * a + d] = nodes[3 * conn[8 * e + a] + d];
                ue[3 * a + d] = U[3 * conn[8 * e + a] + d];
            }
        for (int g = 0; g < 8; ++g) {
            const int64_t gp = 8 * e + g;
            double invJ = 1.0;
            if (kind != 2) {
                const double q = 0.57735026918962576451;
                const double xi[3] = {GSGN[g][0] * q, GSGN[g][1] * q, GSGN[g][2] * q};
                double dNdX[24], dj, F[9];
                or_ref_grads(Xe, xi, dNdX, &dj);
                defgrad(ue, dNdX, F);
                const double J = det3(F);
                invJ = J > 0 ? 1.0 / J : 1.0;
            }
            for (int k = 0; k < 6; ++k) sig[6 * gp + k] = tau[6 * gp + k] * invJ;
        }
    }
}
