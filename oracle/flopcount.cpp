// flopcount.cpp -- instrumented FP64 operation count of the oracle's element routine.
//
// TEST INFRASTRUCTURE ONLY (measurement support): compiles oracle/pfem_oracle.c as C++
// with `double` replaced by a counting type, runs or_element_finite on an elastic and a
// plastic Hex8 F-bar state and prints the floating-point operations per element:
// add/sub/mul/div/fma-free count as 1 each, sqrt/cbrt/log/log1p/exp as 1 each (listed
// separately).  The first count is that of the restatement as written (dense 9x9 A
// times the 9x24 Geff, each column of A from a full unit-dF chain); the second that of
// se_element_finite below, the same element in the structure-exploiting form of the
// device kernels.
//   make -C oracle flopcount   (g++ -O0 -std=c++17 -fpermissive, output oracle/flopcount)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct Counts {
    long long add = 0, mul = 0, div = 0, fn = 0;
};
static Counts g_cnt;

struct fd {
    double v;
    fd() = default;
    fd(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator long long() const { return (long long)v; }
    explicit operator float() const { return (float)v; }
    fd &operator+=(fd o) { ++g_cnt.add; v += o.v; return *this; }
    fd &operator-=(fd o) { ++g_cnt.add; v -= o.v; return *this; }
    fd &operator*=(fd o) { ++g_cnt.mul; v *= o.v; return *this; }
    fd &operator/=(fd o) { ++g_cnt.div; v /= o.v; return *this; }
    fd operator-() const { return fd(-v); }
    fd operator+() const { return *this; }
};
static inline fd operator+(fd a, fd b) { ++g_cnt.add; return fd(a.v + b.v); }
static inline fd operator-(fd a, fd b) { ++g_cnt.add; return fd(a.v - b.v); }
static inline fd operator*(fd a, fd b) { ++g_cnt.mul; return fd(a.v * b.v); }
static inline fd operator/(fd a, fd b) { ++g_cnt.div; return fd(a.v / b.v); }
static inline bool operator<(fd a, fd b) { return a.v < b.v; }
static inline bool operator>(fd a, fd b) { return a.v > b.v; }
static inline bool operator<=(fd a, fd b) { return a.v <= b.v; }
static inline bool operator>=(fd a, fd b) { return a.v >= b.v; }
static inline bool operator==(fd a, fd b) { return a.v == b.v; }
static inline bool operator!=(fd a, fd b) { return a.v != b.v; }
static inline fd fabs(fd a) { return fd(std::fabs(a.v)); }
static inline fd sqrt(fd a) { ++g_cnt.fn; return fd(std::sqrt(a.v)); }
static inline fd cbrt(fd a) { ++g_cnt.fn; return fd(std::cbrt(a.v)); }
static inline fd log(fd a) { ++g_cnt.fn; return fd(std::log(a.v)); }
static inline fd log1p(fd a) { ++g_cnt.fn; return fd(std::log1p(a.v)); }
static inline fd exp(fd a) { ++g_cnt.fn; return fd(std::exp(a.v)); }

#define double fd
#include "pfem_oracle.c"
#undef double

static long long total(const Counts &c) { return c.add + c.mul + c.div + c.fn; }

// Structure-exploiting form of the same element (Hex8 F-bar, isotropic hardening), the
// algorithm of the device kernels with their redundancies taken out: centroid quantities
// once per element; the trial elastic left Cauchy-Green tensor from M = Fb Cpi with the
// symmetric Cpi and only its 6 distinct entries; the consistent tangent D = C - cdev Idev
// + cnn nh nh^T applied in closed form; each column of A = dP/dFb from the rank-2
// perturbation dbe = e_p m^T + m e_p^T (m = column q of M) in the eigenbasis of be and
// the rank-1 change of Fb^-T; the F-bar term as A vec(F) (g0 - gJ)^T; and
// Ke(a,b) = sum_jl dNa_j SA_(ij),(kl) dNb_l contracted in two stages (per-node partial
// sums over j first, then over l), with gJ_b formed once per node b.  main() checks Fe
// and Ke against or_element_finite before printing the count.
static int se_element_finite(const or_mat *m, const fd Xe[24], const fd ue[24], const fd *cpinv,
                             const fd *abar, fd Fe[24], fd Ke[576]) {
    const fd gq = fd(1.0 / std::sqrt(3.0));
    fd dNdX0[24], dj0, F0[9], F0i[9], g0[24];
    const fd z[3] = {fd(0.0), fd(0.0), fd(0.0)};
    or_ref_grads(Xe, z, dNdX0, &dj0);
    defgrad(ue, dNdX0, F0);
    const fd J0 = det3(F0);
    inv3(F0, F0i);
    for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k)
            g0[3 * a + k] = dNdX0[3 * a] * M3(F0i, 0, k) + dNdX0[3 * a + 1] * M3(F0i, 1, k) +
                            dNdX0[3 * a + 2] * M3(F0i, 2, k);
    for (int i = 0; i < 24; ++i) Fe[i] = fd(0.0);
    for (int i = 0; i < 576; ++i) Ke[i] = fd(0.0);
    const int I6[6] = {0, 1, 2, 0, 1, 0}, J6[6] = {0, 1, 2, 1, 2, 2};
    const fd G = fd(m->G), lam = fd(m->lam);
    for (int g = 0; g < 8; ++g) {
        const fd xi3[3] = {fd(GSGN[g][0]) * gq, fd(GSGN[g][1]) * gq, fd(GSGN[g][2]) * gq};
        fd dNdX[24], w, F[9], Fb[9];
        or_ref_grads(Xe, xi3, dNdX, &w);
        defgrad(ue, dNdX, F);
        const fd s = cbrt(J0 / det3(F));
        for (int k = 0; k < 9; ++k) Fb[k] = s * F[k];
        if (det3(Fb) <= fd(0.0)) return 1;
        const fd *c = cpinv + 6 * g;
        fd Mm[9];  // Fb Cpi, Cpi symmetric (00,11,22,01,12,02)
        for (int i = 0; i < 3; ++i) {
            Mm[i] = Fb[i] * c[0] + Fb[i + 3] * c[3] + Fb[i + 6] * c[5];
            Mm[i + 3] = Fb[i] * c[3] + Fb[i + 3] * c[1] + Fb[i + 6] * c[4];
            Mm[i + 6] = Fb[i] * c[5] + Fb[i + 3] * c[4] + Fb[i + 6] * c[2];
        }
        fd be[9];
        for (int t = 0; t < 6; ++t) {
            const fd v = Mm[I6[t]] * Fb[J6[t]] + Mm[I6[t] + 3] * Fb[J6[t] + 3] + Mm[I6[t] + 6] * Fb[J6[t] + 6];
            M3(be, I6[t], J6[t]) = v;
            M3(be, J6[t], I6[t]) = v;
        }
        fd b[3], Q[9], l[3];
        or_eig_sym3(be, b, Q);
        for (int A = 0; A < 3; ++A) l[A] = fd(0.5) * log(b[A]);
        fd ee[6];
        for (int t = 0; t < 6; ++t) {
            const fd v = l[0] * Q[I6[t]] * Q[J6[t]] + l[1] * Q[I6[t] + 3] * Q[J6[t] + 3] +
                         l[2] * Q[I6[t] + 6] * Q[J6[t] + 6];
            ee[t] = t < 3 ? v : fd(2.0) * v;
        }
        // radial return (zero additive plastic strain and back stress, linear hardening)
        const fd tr = ee[0] + ee[1] + ee[2];
        fd tau[6], xi[6], nh[6];
        for (int k = 0; k < 3; ++k) tau[k] = lam * tr + fd(2.0) * G * ee[k];
        for (int k = 3; k < 6; ++k) tau[k] = G * ee[k];
        const fd p = (tau[0] + tau[1] + tau[2]) / fd(3.0);
        for (int k = 0; k < 6; ++k) xi[k] = k < 3 ? tau[k] - p : tau[k];
        const fd nrm = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                            fd(2.0) * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
        const fd s32 = fd(std::sqrt(1.5)), q = s32 * nrm;
        const fd f = q - (fd(m->sy0) + fd(m->Hiso) * abar[g]);
        fd cdev(0.0), cnn(0.0);
        const bool plastic = f > fd(0.0);
        for (int k = 0; k < 6; ++k) nh[k] = fd(0.0);
        if (plastic) {
            for (int k = 0; k < 6; ++k) nh[k] = xi[k] / nrm;
            const fd dg = f / (fd(3.0) * G + fd(m->Hiso));
            const fd r = fd(2.0) * G * dg * s32;
            for (int k = 0; k < 6; ++k) tau[k] = tau[k] - r * nh[k];
            const fd b0 = fd(2.0) * G * dg / q, g0c = fd(2.0) * G / (fd(3.0) * G + fd(m->Hiso));
            cdev = fd(3.0) * G * b0;
            cnn = fd(3.0) * G * (b0 - g0c);
        }
        fd Fbi[9], Fit[9], P[9];
        inv3(Fb, Fbi);
        tr3(Fbi, Fit);
        const fd T[9] = {tau[0], tau[3], tau[5], tau[3], tau[1], tau[4], tau[5], tau[4], tau[2]};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                P[i + 3 * j] = T[i] * Fit[3 * j] + T[i + 3] * Fit[1 + 3 * j] + T[i + 6] * Fit[2 + 3 * j];
        fd gam[3][3];
        for (int A = 0; A < 3; ++A)
            for (int B = 0; B < 3; ++B) {
                if (A == B) gam[A][B] = fd(1.0) / (fd(2.0) * b[A]);
                else if (fabs(b[A] - b[B]) < fd(EIG_TOL) * (fabs(b[A]) + fabs(b[B]) + fd(1.0)))
                    gam[A][B] = fd(1.0) / (b[A] + b[B]);
                else {
                    const fd d = b[A] - b[B];
                    gam[A][B] = fd(0.5) * log1p(d / b[B]) / d;
                }
            }
        const fd sw = s * w;
        fd SA[81], qv[9];
        for (int k = 0; k < 9; ++k) qv[k] = fd(0.0);
        for (int col = 0; col < 9; ++col) {
            const int pp = col % 3, qq = col / 3;
            fd u[3], v[3], Y[3][3], QY[9], de[6], dt[6];
            for (int A = 0; A < 3; ++A) {
                u[A] = Q[pp + 3 * A];
                v[A] = Q[3 * A] * Mm[3 * qq] + Q[1 + 3 * A] * Mm[1 + 3 * qq] + Q[2 + 3 * A] * Mm[2 + 3 * qq];
            }
            for (int A = 0; A < 3; ++A)
                for (int B = A; B < 3; ++B) Y[A][B] = Y[B][A] = gam[A][B] * (u[A] * v[B] + v[A] * u[B]);
            for (int i = 0; i < 3; ++i)
                for (int B = 0; B < 3; ++B) QY[i + 3 * B] = Q[i] * Y[0][B] + Q[i + 3] * Y[1][B] + Q[i + 6] * Y[2][B];
            for (int t = 0; t < 6; ++t) {
                const fd vv = QY[I6[t]] * Q[J6[t]] + QY[I6[t] + 3] * Q[J6[t] + 3] + QY[I6[t] + 6] * Q[J6[t] + 6];
                de[t] = t < 3 ? vv : fd(2.0) * vv;
            }
            const fd trd = de[0] + de[1] + de[2];
            if (plastic) {
                const fd nde = nh[0] * de[0] + nh[1] * de[1] + nh[2] * de[2] + nh[3] * de[3] + nh[4] * de[4] +
                               nh[5] * de[5];
                const fd cd3 = cdev * trd / fd(3.0), cn = cnn * nde;
                const fd ga = fd(2.0) * G - cdev, gs = G - fd(0.5) * cdev, lt = lam * trd + cd3;
                for (int k = 0; k < 3; ++k) dt[k] = lt + ga * de[k] + cn * nh[k];
                for (int k = 3; k < 6; ++k) dt[k] = gs * de[k] + cn * nh[k];
            } else {
                const fd lt = lam * trd, g2 = fd(2.0) * G;
                for (int k = 0; k < 3; ++k) dt[k] = lt + g2 * de[k];
                for (int k = 3; k < 6; ++k) dt[k] = G * de[k];
            }
            const fd dT[9] = {dt[0], dt[3], dt[5], dt[3], dt[1], dt[4], dt[5], dt[4], dt[2]};
            fd wq[3];
            for (int r = 0; r < 3; ++r) wq[r] = T[r] * Fit[3 * qq] + T[r + 3] * Fit[1 + 3 * qq] + T[r + 6] * Fit[2 + 3 * qq];
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) {
                    const fd dP = dT[r] * Fit[3 * cc] + dT[r + 3] * Fit[1 + 3 * cc] + dT[r + 6] * Fit[2 + 3 * cc] -
                                  wq[r] * Fit[pp + 3 * cc];
                    SA[(r + 3 * cc) + 9 * col] = sw * dP;
                    qv[r + 3 * cc] += dP * F[col];
                }
        }
        fd Fi[9], gd[24], wP[9];
        inv3(F, Fi);
        const fd sw3 = sw / fd(3.0);
        for (int k = 0; k < 9; ++k) { qv[k] = sw3 * qv[k]; wP[k] = w * P[k]; }
        for (int bn = 0; bn < 8; ++bn)  // g0 - gJ, once per node
            for (int k = 0; k < 3; ++k)
                gd[3 * bn + k] = g0[3 * bn + k] - (dNdX[3 * bn] * M3(Fi, 0, k) + dNdX[3 * bn + 1] * M3(Fi, 1, k) +
                                                    dNdX[3 * bn + 2] * M3(Fi, 2, k));
        for (int a = 0; a < 8; ++a) {
            const fd *da = dNdX + 3 * a;
            fd Ta[3][9], h[3];
            for (int cc = 0; cc < 9; ++cc)
                for (int i = 0; i < 3; ++i)
                    Ta[i][cc] = da[0] * SA[i + 9 * cc] + da[1] * SA[i + 3 + 9 * cc] + da[2] * SA[i + 6 + 9 * cc];
            for (int i = 0; i < 3; ++i) {
                h[i] = da[0] * qv[i] + da[1] * qv[i + 3] + da[2] * qv[i + 6];
                Fe[3 * a + i] += da[0] * wP[i] + da[1] * wP[i + 3] + da[2] * wP[i + 6];
            }
            for (int bn = 0; bn < 8; ++bn) {
                const fd *db = dNdX + 3 * bn;
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k)
                        Ke[(3 * a + i) + 24 * (3 * bn + k)] +=
                            Ta[i][k] * db[0] + Ta[i][k + 3] * db[1] + Ta[i][k + 6] * db[2] + h[i] * gd[3 * bn + k];
            }
        }
    }
    return 0;
}

static double max_rel_diff(const fd *x, const fd *y, int n) {
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
        num = std::max(num, std::fabs(x[i].v - y[i].v));
        den = std::max(den, std::fabs(y[i].v));
    }
    return den > 0 ? num / den : num;
}

int main() {
    or_mat m;
    if (or_material_init(&m, 200e9, 0.3, 250e6, 1e9, 0.0, 250e6, 0.0)) return 1;
    // a slightly distorted unit-cube element (h = 0.01)
    fd Xe[24], ue[24], ep[48], bt[48], ab[8], cp[48], sg[48], Fe[24], Ke[576];
    const int nat[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int a = 0; a < 8; ++a)
        for (int d = 0; d < 3; ++d) Xe[3 * a + d] = fd(0.01 * nat[a][d] + 0.0007 * ((a * 3 + d * 5) % 7 - 3) / 3.0);
    const char *label[2] = {"elastic", "plastic"};
    const double amp[2] = {1e-6, 5e-4};  // nodal displacement amplitudes (m)
    for (int k = 0; k < 2; ++k) {
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d)
                ue[3 * a + d] = fd(amp[k] * ((a * 7 + d * 3) % 5 - 2) + (d == 0 ? -amp[k] * 4 * nat[a][0] : 0.0));
        for (int i = 0; i < 48; ++i) { ep[i] = fd(0.0); bt[i] = fd(0.0); sg[i] = fd(0.0); cp[i] = fd(i % 6 < 3 ? 1.0 : 0.0); }
        for (int i = 0; i < 8; ++i) ab[i] = fd(0.0);
        int g = -1;
        fd J(0.0);
        g_cnt = Counts{};
        int st = or_element_finite(1, &m, Xe, ue, ep, bt, ab, cp, sg, 0, Fe, Ke, &g, &J);
        const Counts c = g_cnt;
        int nplastic = 0;
        // a second pass with commit to read the plastic multiplier increments
        fd ab2[8];
        for (int i = 0; i < 8; ++i) ab2[i] = ab[i];
        or_element_finite(1, &m, Xe, ue, ep, bt, ab2, cp, sg, 1, Fe, Ke, &g, &J);
        for (int i = 0; i < 8; ++i) nplastic += ab2[i].v > 0;
        printf("%s: status %d, plastic Gauss points %d/8: flops per element %lld "
               "(add/sub %lld, mul %lld, div %lld, sqrt/cbrt/log/exp %lld)\n",
               label[k], st, nplastic, total(c), c.add, c.mul, c.div, c.fn);
        // the structure-exploiting form on the same state, checked against the oracle's Fe, Ke
        fd Fe0[24], Ke0[576], Fe1[24], Ke1[576];
        for (int i = 0; i < 48; ++i) { ep[i] = fd(0.0); bt[i] = fd(0.0); cp[i] = fd(i % 6 < 3 ? 1.0 : 0.0); }
        for (int i = 0; i < 8; ++i) ab[i] = fd(0.0);
        or_element_finite(1, &m, Xe, ue, ep, bt, ab, cp, sg, 0, Fe0, Ke0, &g, &J);
        g_cnt = Counts{};
        const int st2 = se_element_finite(&m, Xe, ue, cp, ab, Fe1, Ke1);
        const Counts c2 = g_cnt;
        printf("%s_structured: status %d, plastic Gauss points %d/8: flops per element %lld "
               "(add/sub %lld, mul %lld, div %lld, sqrt/cbrt/log/exp %lld); max rel diff Fe %.3e Ke %.3e\n",
               label[k], st2, nplastic, total(c2), c2.add, c2.mul, c2.div, c2.fn, max_rel_diff(Fe1, Fe0, 24),
               max_rel_diff(Ke1, Ke0, 576));
    }
    return 0;
}
