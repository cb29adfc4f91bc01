// flopcount.cpp -- instrumented FP64 operation count of the oracle's element routine.
//
// TEST INFRASTRUCTURE ONLY (measurement support): compiles oracle/pfem_oracle.c as C++
// with `double` replaced by a counting type, runs or_element_finite on an elastic and a
// plastic Hex8 F-bar state and prints the floating-point operations per element:
// add/sub/mul/div/fma-free count as 1 each, sqrt/cbrt/log/log1p/exp as 1 each (listed
// separately).  The count is that of the restatement as written (dense 9x9 A times the
// 9x24 Geff, structure-exploiting G^T contractions).
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
    }
    return 0;
}
