// cad_wave_theory.hip -- the wave-maker's theory functions: MODEL WAVE-BC
// <side> FUNC STOKES (fifth order), CNOIDAL (third order), STK-CND (one of
// the two by the Ursell number) and STREAM n (the stream function of order
// n, vf_wsfmb2.f), applied in VF_BWUWN (vf_bwuwn.f:257-480) through
// VF_CWMAK0/1/2 (vf_cwmak0.f, vf_cwmak1.f, vf_cwmak2.f).
//
// The reference evaluates these waves in single precision, in CGS units, with
// the C library's tanhf, coshf, sinhf, expf and sincosf (vf_wstk0.f,
// vf_wcnd0.f).  The restatement below runs on the host, in float, calling the
// same library, so each face velocity carries the reference's bits.  Each
// call of the boundary pass then goes:
//   host    the side's wave (VF_CWMAK0), phase and ramp, the level at the
//           phase (VF_CWMAK1);
//   device  one thread per column: the column's water level and the depth
//           scale VWS = (level + depth) / (measured level + depth);
//   host    VF_CWMAK2 at every face height of every column,
//           ZC = VWS (z - W-LEVEL + depth) - depth;
//   device  one thread per column writes the side's boundary faces.
// The host reads the scales back and the device waits for the velocities, one
// round trip per side per boundary pass: the work is O(columns x cells up a
// column) on a face of the domain.
//
// VF_STK0 keeps its last secant step between calls (SAVE) and starts the next
// call's first step with its sign; the context keeps it (wv_stk_dddk) and the
// calls follow the reference's: one at set-up (VF_CSETUP, vf_csetup.f:113),
// then one per side per boundary pass.
#include "cad_common.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

#define RET_IF_W(x)            \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

namespace {

constexpr double PI_REF = 3.141592653589794;
constexpr double ZERO_REF = 1.0e-20;

// constants of the wave routines (single precision, CGS)
constexpr float G_CGS = 980.0f;
constexpr float PI_S = 3.141593f;

// c[0] + (c[1] + (c[2] + ... )*r)*r, innermost first: the nested polynomial
// form of the coefficient formulas, in their order of operations
template <int N>
float nest(float r, const float (&c)[N])
{
    float acc = c[N - 1];
    for (int i = N - 2; i >= 0; i--) acc = c[i] + acc * r;
    return acc;
}

inline float sq(float x) { return x * x; }
// X**4 with a constant exponent: a left-to-right product, as the reference's
// compiler evaluates it (squaring twice rounds differently and moves the
// cnoidal wave's level by a few ulps)
inline float p4(float x) { return x * x * x * x; }

// ---- fifth-order Stokes wave (vf_wstk0.f) --------------------------------

struct Stokes {
    int n;
    float rk, dk, cc, c0, b0, rl;
    float a[5], b[5];
};

// VF_HK (vf_wstk0.f): kh of the linear dispersion relation kh tanh kh = ww,
// Newton from sqrt(ww), at most 10 steps, |dk/k| < 1e-6
float vf_hk(float ww)
{
    float hk = ww;
    if (ww > 15.0f) return hk;
    hk = std::sqrt(ww);
    for (int n = 1; n <= 10; n++) {
        const float t = tanhf(hk);
        const float d = (hk * t - ww) / (t + hk * (1.0f - t * t));
        hk = hk - d;
        if (std::fabs(d / hk) < 1.0e-6f) return hk;
    }
    return hk;
}

// the wave-speed factor C1 of order n at kh = dk1 (vf_wstk0.f, labels 12-14)
float stk_c1(int n, float dk1, float hd, float *th_out)
{
    const float th = tanhf(dk1);
    const float ct = 1.0f / th;
    const float r = ct * ct;
    const float f = sq(hd * dk1);
    float c1 = 0.0f;
    if (n >= 4)
        c1 = nest(r, {39.0f, 351.0f, -2194.0f, 2454.0f, -117.0f, -405.0f}) / 1024.0f -
             nest(r, {3.0f, -7.0f, -3.0f, -9.0f}) * ct / (64.0f * dk1);
    if (n >= 2) c1 = c1 * f + nest(r, {9.0f, -10.0f, 9.0f}) / 16.0f - ct / (2.0f * dk1);
    c1 = c1 * f + 1.0f;
    *th_out = th;
    return c1;
}

// VF_STK0: kh by secant steps on sqrt(kh tanh kh) C1 = sigma sqrt(d/g), then
// the coefficients A (surface), B (velocity) and B0.  d, t, h in cm and s.
void stk0(Stokes &s, float &dddk, int n, float d, float t, float h)
{
    const float eps = 1.0e-5f;
    s.n = n;
    const float sgm = 2.0f * PI_S / t;
    const float ww = sgm * sgm * d / G_CGS;
    const float w = std::sqrt(ww);
    const float hd = h / (2.0f * d);
    float dk = vf_hk(ww);
    float th = 0.0f, c1 = 0.0f, r1 = 0.0f, r2 = 0.0f, ddk = 0.0f;
    bool done = false;
    for (int iter = 1; iter <= 20 && !done; iter++) {
        ddk = std::copysign(dk * eps, dddk);
        float dk1 = dk + ddk;
        c1 = stk_c1(n, dk1, hd, &th);
        r1 = std::sqrt(dk1 * th) * c1 - w;
        if (dk1 != dk) {
            r2 = r1;
            dk1 = dk;
            c1 = stk_c1(n, dk1, hd, &th);
            r1 = std::sqrt(dk1 * th) * c1 - w;
        }
        dddk = -r1 * ddk / (r2 - r1);
        dk = dk + dddk;
        if (std::fabs(dddk / dk) >= eps) continue;
        // converged: the factor at the final kh
        c1 = stk_c1(n, dk, hd, &th);
        done = true;
    }
    if (!done) dk = dk - dddk;
    const float rk = dk / d;
    s.rl = 2.0f * PI_S / rk;
    const float e = dk * hd;
    const float c0 = std::sqrt(G_CGS * th / rk);
    s.rk = rk;
    s.dk = dk;
    s.c0 = c0;
    s.cc = c1 * c0;
    const float ct = 1.0f / th;
    const float r = ct * ct;
    const float e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e;
    float *a = s.a, *b = s.b;
    for (int i = 0; i < 5; i++) a[i] = b[i] = 0.0f;
    float b0 = 0.0f;
    const float r5a = nest(r, {1.0f, 5.0f}), r5b = nest(r, {3.0f, 5.0f});
    if (n >= 5) {
        a[0] = e5 * nest(r, {-225.0f, 2370.0f, 14622.0f, -6070.0f, 53932.0f, -340410.0f, -764046.0f,
                             203310.0f, 188325.0f}) /
               (12288.0f * r5a * r5b);
        a[2] = e5 * nest(r, {-180.0f, -1008.0f, 900.0f, -5328.0f, 46980.0f, -7776.0f, -14580.0f}) * r /
               (4096.0f * r5a);
        a[4] = e5 * nest(r, {45.0f, -150.0f, -570.0f, 2618.0f, -3896.0f, -498.0f, 25866.0f, 8910.0f,
                             6075.0f}) *
               5.0f / (12288.0f * r5a * r5b);
        b[0] = e5 * nest(r, {861.0f, 12262.0f, 30114.0f, -75986.0f, -131460.0f, -91438.0f, -199602.0f,
                             89370.0f, 62775.0f}) /
               (4096.0f * r5a * r5b);
        b[2] = e5 * nest(r, {173.0f, 1413.0f, 1665.0f, -6591.0f, -7569.0f, 10935.0f, 4131.0f, -3645.0f}) /
               (1024.0f * r5a);
        b[4] = e5 * nest(r, {1083.0f, 15230.0f, 39970.0f, -78730.0f, -89928.0f, 157690.0f, -16290.0f,
                             -39150.0f, 10125.0f}) /
               (4096.0f * r5a * r5b);
    }
    if (n >= 4) {
        a[1] = e4 * nest(r, {12.0f, -352.0f, 792.0f, 0.0f, -324.0f}) * ct / 384.0f;
        a[3] = e4 * nest(r, {21.0f, 1.0f, -262.0f, 522.0f, 81.0f, 405.0f}) * ct / (384.0f * r5a);
        b[1] = e4 * nest(r, {-103.0f, 19.0f, 261.0f, 81.0f, -162.0f}) * ct / 192.0f;
        b[3] = e4 * nest(r, {-197.0f, -747.0f, 1466.0f, 342.0f, -1269.0f, 405.0f}) * ct / (384.0f * r5a);
        b0 = e4 * nest(r, {-3.0f, 7.0f, 3.0f, 9.0f}) * ct / (64.0f * dk);
    }
    if (n >= 3) {
        a[0] = a[0] + e3 * nest(r, {1.0f, -3.0f, 3.0f, -9.0f}) * 3.0f / 64.0f;
        a[2] = a[2] + e3 * nest(r, {-1.0f, 3.0f, -3.0f, 9.0f}) * 3.0f / 64.0f;
        b[0] = b[0] + e3 * nest(r, {39.0f, -41.0f, -3.0f, -27.0f}) / 64.0f;
        b[2] = b[2] + e3 * nest(r, {13.0f, 17.0f, -57.0f, 27.0f}) / 64.0f;
    }
    if (n >= 2) {
        a[1] = a[1] + e2 * nest(r, {-1.0f, 3.0f}) * ct / 4.0f;
        b[1] = b[1] + e2 * nest(r, {-1.0f, 1.0f}) * ct * 3.0f / 4.0f;
        b0 = b0 + e2 * (-ct / (2.0f * dk));
    }
    a[0] = a[0] + e;
    b[0] = b[0] + e;
    for (int i = 1; i <= n; i++) {
        a[i - 1] = a[i - 1] / rk;
        b[i - 1] = b[i - 1] * c0 * (float)i / sinhf((float)i * dk);
    }
    s.b0 = b0 * c0;
}

// VF_STK1: the surface elevation (cm) at phase x
float stk1(const Stokes &s, float x)
{
    const float xk = 2.0f * PI_S * x;
    float zs = 0.0f;
    for (int i = 1; i <= s.n; i++) {
        float si, co;
        sincosf(xk * (float)i, &si, &co);
        zs = zs + s.a[i - 1] * co;
    }
    return zs;
}

// VF_STK2: horizontal and vertical velocity (cm/s) at phase x, height z (cm,
// from the still water level)
void stk2(const Stokes &s, float x, float z, float *u, float *w)
{
    const float xk = 2.0f * PI_S * x;
    const float zk = z * s.rk + s.dk;
    float uu = s.b0, ww = 0.0f;
    for (int i = 1; i <= s.n; i++) {
        float si, co;
        sincosf(xk * (float)i, &si, &co);
        const float zi = zk * (float)i;
        const float ch = coshf(zi), sh = sinhf(zi);
        uu = uu + s.b[i - 1] * ch * co;
        ww = ww + s.b[i - 1] * sh * si;
    }
    *u = uu;
    *w = ww;
}

// ---- third-order cnoidal wave (vf_wcnd0.f) -------------------------------

struct Cnoidal {
    int n;
    float d, ql, rl;
    float a[4], b[3][4], bb[3][4];   // b[j][i]: B(J,I)
};

// VF_ELFN (vf_wcnd0.f): the Jacobi elliptic functions cn, sn, dn at x
// (periods of 1) for the nome exponent ql, by the theta series in the form
// the reference picks from IQ = int((12 ql/pi - 7)/5)
void elfn(float ql, float x, float *cn, float *sn, float *dn)
{
    const float pi = 3.141593f, pisq = 9.869604f, pi2 = 1.570796f;
    const int iq = (int)((12.0f * ql / pi - 7.0f) / 5.0f);
    const float qql = pisq / ql;
    const float q = expf(-ql);
    const float qq = expf(-qql);
    float q2 = 0, q4 = 0, q6 = 0, rkk = 0, p = 0, t02 = 0, t03 = 0, t04 = 0, fc = 0, fs = 0, fd = 0;
    if (iq < 0) {
        q2 = 2.0f * qq;
        rkk = 4.0f * std::sqrt(qq);
        p = 2.0f * qql;
    } else if (iq == 0) {
        q2 = qq * qq;
        q4 = q2 * q2;
        q6 = q4 * q2;
        t02 = 1.0f + 2.0f * (-qq + q4);
        t03 = 1.0f + 2.0f * (qq + q4);
        t04 = 1.0f + q2 + q6;
        rkk = 4.0f * std::sqrt(qq) * sq(t04 / t03);
        p = 2.0f * qql;
    } else {
        q2 = q * q;
        q4 = q2 * q2;
        q6 = q4 * q2;
        t02 = 1.0f + q2 + q6;
        t03 = 1.0f + 2.0f * (q + q4);
        t04 = 1.0f + 2.0f * (-q + q4);
        rkk = sq(t04 / t03);
        p = 2.0f * pi;
    }
    if (iq >= 0) {
        fc = t04 / t02;
        fs = t03 / t02;
        fd = t04 / t03;
    }
    if (std::fmod(x, 0.25f) == 0.0f) {
        // quarter periods exactly (label 450)
        const int m = (int)(x / 0.25f);
        if (m % 2 == 0) {
            *cn = (m % 4 != 0) ? -1.0f : 1.0f;
            *sn = 0.0f;
            *dn = 1.0f;
        } else {
            *cn = 0.0f;
            *sn = ((m - 1) % 4 != 0) ? -1.0f : 1.0f;
            *dn = rkk;
        }
        return;
    }
    float x1 = std::fmod(x, 1.0f);
    if (std::fabs(x1) > 0.5f) x1 = x1 - std::copysign(1.0f, x1);
    const float sig = std::copysign(1.0f, 0.25f - std::fabs(x1));
    if (sig < 0.0f) x1 = x1 - std::copysign(0.5f, x1);
    if (iq <= 0) {
        const float xx = p * x1;
        const float c = coshf(xx), s = sinhf(xx);
        if (iq != 0) {
            const float t = s / c;
            const float c2 = q2 * (2.0f * c * c - 1.0f);
            *cn = (1.0f - c2) / c * sig;
            *sn = t * sig;
            *dn = (1.0f + c2) / c;
            return;
        }
        const float c2 = 4.0f * c * c;
        const float s2 = 4.0f * s * s;
        const float t4 = c * (t04 + s2 * (q2 + q6 * (c2 - 1.0f)));
        *cn = fc / t4 * (t03 + c2 * (-qq + q4 * s2)) * sig;
        *sn = fs / t4 * s * (t04 + c2 * (-q2 + q6 * (s2 + 1.0f))) * sig;
        *dn = fd / t4 * (t02 + c2 * (qq + q4 * s2));
        return;
    }
    const float xx = p * x1;
    float s, c;
    sincosf(xx, &s, &c);
    const float c2 = 4.0f * c * c;
    const float s2 = 4.0f * s * s;
    const float t4 = t03 - c2 * (q + q4 * s2);
    *cn = fc / t4 * c * (t02 - s2 * (q2 + q6 * (c2 - 1.0f))) * sig;
    *sn = fs / t4 * s * (t02 - c2 * (q2 + q6 * (s2 - 1.0f))) * sig;
    *dn = fd / t4 * (t04 + c2 * (q - q4 * s2));
}

// VF_CND0: the nome from the Ursell number by fixed-point steps (at most 20,
// |dq/q| < 1e-5; Ursell numbers below 10 are taken as 10), then the
// coefficients of the surface (A) and of the velocity (B, BB) to order n.
void cnd0(Cnoidal &s, int n, float d, float t, float h)
{
    const float eps = 1.0e-5f;
    s.n = n;
    s.d = d;
    const int nn = n + 1;
    const float gd = G_CGS * d;
    const float c0 = std::sqrt(gd);
    const float e = h / d;
    const float e2 = e * e;
    const float e3 = e2 * e;
    float ur = G_CGS * h * t * t / (d * d);
    if (ur < 10.0f) ur = 10.0f;
    const float ur34 = ur * 0.75f;
    float qql = std::sqrt(ur34);
    float c00 = 0, c01 = 0, c10 = 0, c11 = 0, c20 = 0;
    if (n >= 3) {
        c00 = e2 * 0.2f;
        c01 = e2 * (-0.25f);
        c10 = e2 * 0.825f;
        c11 = e2 * (-0.5f);
        c20 = e2 * 0.825f;
    }
    if (n >= 2) {
        c00 = c00 + e * (-0.25f);
        c10 = c10 + e * (-0.5f);
    }
    c00 = c00 + 1.0f;
    float r = 0, o = 0, t3 = 0;
    for (int i = 1; i <= 20; i++) {
        const float qq = expf(-qql);
        const float q2 = qq * qq;
        const float q4 = q2 * q2;
        const float t2 = p4(1.0f + 2.0f * (-qq + q4));
        t3 = p4(1.0f + 2.0f * (qq + q4));
        const float t4 = p4(1.0f + q2 * (1.0f + q4));
        const float sv = 1.0f + 8.0f * q2 * (1.0f - q2);
        r = 16.0f * qq * t4 / t2;
        o = (2.0f / qql + t3 - sv) / t2;
        const float qqll = qql;
        qql = std::sqrt(ur34 * (c00 + r * (c10 + r * c20) + o * (c01 + r * c11)) / t2);
        if (std::fabs((qqll - qql) / qql) < eps) break;
    }
    float *a = s.a;
    for (int i = 0; i < nn; i++) {
        a[i] = 0.0f;
        for (int j = 0; j < n; j++) s.b[j][i] = 0.0f;
    }
    float c = 0.0f;
    const float rr = r * r, ro = r * o, oo = o * o;
    const float rrr = rr * r, rro = rr * o, roo = ro * o, ooo = oo * o;
    auto B = [&](int j, int i) -> float & { return s.b[j - 1][i - 1]; };
    if (n >= 3) {
        a[0] = e3 * (133.0f * r - 16.0f * o + 399.0f * rr - 466.0f * ro + 100.0f * oo + 266.0f * rrr -
                     466.0f * rro + 200.0f * roo) /
               400.0f;
        a[1] = e3 * (50.0f - r - 60.0f * o) / 80.0f;
        a[2] = e3 * (-151.0f + r + 60.0f * o) / 80.0f;
        a[3] = e3 * 101.0f / 80.0f;
        B(1, 1) = e3 * (-71.0f * r + 47.0f * o - 23.0f * rr + 97.0f * ro - 50.0f * oo + 153.0f * rrr -
                        153.0f * rro - 25.0f * roo + 25.0f * ooo) /
                  200.0f;
        B(2, 1) = e3 * (6.0f * r + 24.0f * rr - 21.0f * ro) / 8.0f;
        B(3, 1) = e3 * (3.0f * r - 3.0f * rr) / 16.0f;
        B(1, 2) = e3 * (-19.0f - 27.0f * r + 10.0f * o + 101.0f * rr - 100.0f * ro + 15.0f * oo) / 40.0f;
        B(2, 2) = e3 * (6.0f + 36.0f * r - 21.0f * o - 24.0f * rr + 21.0f * ro) / 4.0f;
        B(3, 2) = e3 * (6.0f - 39.0f * r + 6.0f * rr) / 16.0f;
        B(1, 3) = e3 * (-2.0f + 32.0f * r - 15.0f * o) / 10.0f;
        B(2, 3) = e3 * (30.0f - 120.0f * r + 63.0f * o) / 8.0f;
        B(3, 3) = e3 * (-45.0f + 45.0f * r) / 16.0f;
        B(1, 4) = e3 * 6.0f / 5.0f;
        B(2, 4) = e3 * (-15.0f) / 2.0f;
        B(3, 4) = e3 * 45.0f / 16.0f;
        c = e3 * (150.0f + 1079.0f * r - 203.0f * o + 2337.0f * rr - 2653.0f * ro + 350.0f * oo +
                  1558.0f * rrr - 2653.0f * rro + 700.0f * roo + 175.0f * ooo) /
            2800.0f;
    }
    if (n >= 2) {
        a[0] = a[0] + e2 * (-2.0f * r + o - 2.0f * rr + 2.0f * ro) / 4.0f;
        a[1] = a[1] + e2 * (-0.75f);
        a[2] = a[2] + e2 * (0.75f);
        B(1, 1) = B(1, 1) + e2 * (r - o - 2.0f * rr + 2.0f * oo) / 4.0f;
        B(2, 1) = B(2, 1) + e2 * (-0.75f * r);
        B(1, 2) = B(1, 2) + e2 * (1.0f - 6.0f * r + 2.0f * o) / 4.0f;
        B(2, 2) = B(2, 2) + e2 * (-3.0f + 3.0f * r) / 2.0f;
        B(1, 3) = B(1, 3) + e2 * (-1.0f);
        B(2, 3) = B(2, 3) + e2 * 2.25f;
        c = c + e2 * (-6.0f - 16.0f * r + 5.0f * o - 16.0f * rr + 10.0f * ro + 15.0f * oo) / 40.0f;
    }
    a[0] = a[0] + e * (r - o);
    a[1] = a[1] + e;
    B(1, 1) = B(1, 1) + e * (r - o);
    B(1, 2) = B(1, 2) + e;
    c = c + e * (1.0f + 2.0f * r - 3.0f * o) / 2.0f + 1.0f;
    const float cc = c * c0;
    const float rl = t * cc;
    s.rl = rl;
    const float dl = d / rl;
    const float sk = 2.0f * std::sqrt(t3) * qql * dl;
    for (int i = 1; i <= nn; i++) {
        a[i - 1] = a[i - 1] * d;
        for (int j = 1; j <= n; j++) {
            B(j, i) = B(j, i) * c0;
            s.bb[j - 1][i - 1] = B(j, i) * sk * (float)(i - 1) / (float)(2 * j - 1);
        }
    }
    s.ql = PI_S * PI_S / qql;
}

// VF_CND1: the surface elevation (cm) at phase x
float cnd1(const Cnoidal &s, float x)
{
    float cn, sn, dn;
    elfn(s.ql, x / 2.0f, &cn, &sn, &dn);
    const float cn2 = cn * cn;
    const int nn = s.n + 1;
    float zs = 0.0f;
    for (int i = 1; i <= nn; i++) zs = zs * cn2 + s.a[nn - i];
    return zs;
}

// VF_CND2: horizontal and vertical velocity (cm/s) at phase x, height z (cm)
void cnd2(const Cnoidal &s, float x, float z1, float *u1, float *w1)
{
    float cn, sn, dn;
    elfn(s.ql, x / 2.0f, &cn, &sn, &dn);
    const float cn2 = cn * cn;
    const float csd = cn * sn * dn;
    const int n = s.n, nn = n + 1;
    const float z = 1.0f + z1 / s.d;
    const float zz = z * z;
    float u = 0.0f, w = 0.0f;
    for (int i = 1; i <= nn; i++) {
        const int k = nn + 1 - i;
        float uuu = 0.0f, www = 0.0f;
        for (int j = 1; j <= n; j++) {
            const int l = nn - j;
            uuu = uuu * zz + s.b[l - 1][k - 1];
            www = www * zz + s.bb[l - 1][k - 1];
        }
        u = u * cn2 + uuu;
        if (i != nn) w = w * cn2 + www;
    }
    *u1 = u;
    *w1 = w * csd * z;
}

// ---- the stream-function wave (vf_wsfmb2.f) ------------------------------
// Dean's stream-function method B with the second definition of the wave
// speed: psi = (C - U) z - sum_n A_n sinh(n k (d + z)) cos(n k x) / (n k),
// the coefficients A_1..A_N, the length L and psi at the surface fitted by
// Newton least squares to the free-surface condition at 51 phases over half
// a wave, the Bernoulli constant Q adjusted by halving steps until the
// height comes out as asked.  The reference computes in double precision
// from single-precision inputs (SI units, g = 9.8) and keeps its state in
// COMMON blocks; VF_CWMAK0 starts it afresh each call (vf_prep runs again
// and psi_eta is reset, vf_wsfmb2.f:43-50), so one Stream per VF_CWMAK0 call
// holds the same values.  Literals written in single precision in the
// reference keep their single-precision values here ((double)0.3f, ...).
struct Stream {
    // /vf_coef/
    double d, t, h, u, a[24], rk, sigma, b, pseta;
    int nn, nn1, nn2;
    // /vf_cnst/, /vf_sico/
    double pi, pi2, g, si[201], co[201];
    int imax, imax1;
    // /vf_parm/
    double epsa, depsa[24], epsh, depsh;
    int iteram, iterhm, irtry;
    // /vf_euwt/: the phase and height caches of VF_SFMB1/2 are single precision
    double sit, cot, sh, ch, eta, epse, depse;
    float xl, xl2, z;
    int iterem, iecon;
    // /vf_iniv/, /vf_conv/, /vf_derv/, /vf_atmp/
    double rkd0, rk0, rl0, b0, q0, deltq0, depsq, bq, hh;
    int iconv, itera;
    double e1, ekl[23], ek[23 * 23], detaa[23], a2[24];
    // VF_SFMB1/2 locals read on their cached paths before being set there
    // (kept from the previous call, as storage that outlives the call gives)
    double ax1, uu1, ww1, uut1, wwt1, pp1;
    double &EK(int k, int l) { return ek[(k - 1) + 23 * (l - 1)]; }   // EK(k,l), column-major
};

// vf_prep (vf_wsfmb2.f:224-277): constants and the 51 phases' sines
void sf_prep(Stream &s)
{
    s.imax = 50;
    s.epsa = (double)1.0e-3f;
    s.epsh = (double)1.0e-3f;
    s.epse = (double)1.0e-5f;
    s.iteram = 50;
    s.iterhm = 200;
    s.iterem = 30;
    s.irtry = 5;
    s.pi = 3.14159265358979;
    s.pi2 = 2.0 * s.pi;
    s.g = 9.8;
    const double df = s.pi / (double)(float)s.imax;
    s.imax1 = s.imax + 1;
    for (int i = 1; i <= s.imax1; i++) {
        const double f = df * (double)(float)(i - 1);
        s.si[i - 1] = std::sin(f);
        s.co[i - 1] = std::cos(f);
    }
}

// vf_wcdspr (vf_wsfmb2.f:344-362): kd of the linear dispersion relation with
// a current, by Newton
double sf_wcdspr(double w2, double usd)
{
    const double epsk = (double)1.0e-6f;
    const double uw2 = 2.0 * usd * w2;
    const double w = std::sqrt(w2);
    double rkd = w / (1.0 + w * usd);
    for (int i = 1; i <= 10; i++) {
        const double th = std::tanh(rkd);
        const double q = 1.0 - rkd * usd;
        const double f = rkd * th - w2 * (q * q);
        const double df = th + rkd * (1.0 - th * th) + uw2 * (1.0 - rkd * usd);
        const double dkd = f / df;
        rkd = rkd - dkd;
        if (std::fabs(dkd / rkd) < epsk) return rkd;
    }
    return rkd;
}

// vf_init (vf_wsfmb2.f:279-342): first guesses and tolerances
void sf_init(Stream &s)
{
    s.sigma = s.pi2 / s.t;
    const double w2 = s.sigma * s.sigma * s.d / s.g;
    const double usd = (s.u + s.pseta / s.d) / (s.sigma * s.d);
    s.rkd0 = sf_wcdspr(w2, usd);
    s.rk0 = s.rkd0 / s.d;
    s.rl0 = s.pi2 / s.rk0;
    s.b0 = s.rl0 / s.t - (s.u + s.pseta / s.d);
    s.q0 = s.b0 * s.b0 / 2.0;
    s.deltq0 = s.q0 * (double)0.01f;
    s.depsq = s.h * s.h * s.rk0 / (8.0 * std::sinh(2.0 * s.rkd0)) * s.epsh;
    s.depse = s.epse * s.d;
    s.nn1 = s.nn + 1;
    s.nn2 = s.nn + 2;
    const double depsa1 = s.epsa * s.b0 * s.h / 2.0;
    const double sh1 = std::sinh(s.rkd0), ch1 = std::cosh(s.rkd0);
    double shn = sh1, chn = ch1;
    for (int n = 1; n <= s.nn; n++) {
        s.depsa[n - 1] = depsa1 / ((double)(float)n * shn);
        const double c = chn * ch1 + shn * sh1;
        shn = shn * ch1 + chn * sh1;
        chn = c;
    }
    s.depsa[s.nn1 - 1] = s.epsa * s.rl0;
    s.depsa[s.nn2 - 1] = s.epsa * s.b0 * s.h / 2.0;
    s.depsh = s.epsh * s.h;
}

// vf_inita (vf_wsfmb2.f:468-493): starting coefficients
void sf_inita(Stream &s)
{
    const double hinit = std::fmin((double)0.142f * s.rl0, (double)0.82f * s.d) * (double)0.1f;
    s.a[0] = -(hinit * s.b0) / (2.0 * std::sinh(s.rkd0));
    for (int n = 2; n <= s.nn; n++) s.a[n - 1] = 0.0;
    s.a[s.nn1 - 1] = s.rl0;
    s.a[s.nn2 - 1] = 0.0;
    for (int k = 1; k <= s.nn1; k++) s.detaa[k - 1] = 0.0;
}

// vf_fact (vf_wsfmb2.f:495-512): pull the coefficients back towards the start
void sf_fact(Stream &s)
{
    const double fac = (double)0.8f;
    for (int n = 1; n <= s.nn; n++) s.a[n - 1] = s.a[n - 1] * fac;
    s.a[s.nn1 - 1] = (s.a[s.nn1 - 1] - s.rl0) * fac + s.rl0;
    s.a[s.nn2 - 1] = s.a[s.nn2 - 1] * fac;
}

// vf_caleta (vf_wsfmb2.f:745-800): the surface eta at the phase (sit, cot),
// Newton from the current eta; iecon = 1 (eta = 0) if it does not converge
void sf_caleta(Stream &s)
{
    const double rkd = s.rk * s.d;
    s.iecon = 0;
    for (int j = 1; j <= s.iterem; j++) {
        double sitn = s.sit, cotn = s.cot;
        const double rketa = s.rk * s.eta;
        if (std::fabs(rketa) > 2.0) break;
        const double rkdeta = rkd + rketa;
        const double sh1 = std::sinh(rkdeta), ch1 = std::cosh(rkdeta);
        double shn = sh1, chn = ch1;
        double shc = s.b * s.eta - s.pseta;
        double chc = s.b;
        for (int n = 1; n <= s.nn; n++) {
            const double ac = s.a[n - 1] * cotn;
            shc = shc + ac * shn;
            chc = chc + (double)(float)n * s.rk * ac * chn;
            double c = cotn * s.cot - sitn * s.sit;
            sitn = sitn * s.cot + cotn * s.sit;
            cotn = c;
            c = chn * ch1 + shn * sh1;
            shn = shn * ch1 + chn * sh1;
            chn = c;
        }
        const double dd = shc / chc;
        s.eta = s.eta - dd;
        if (std::fabs(dd) < s.depse) return;
    }
    s.iecon = 1;
    s.eta = 0.0;
}

// vf_psieta (vf_wsfmb2.f:666-743): psi at the surface (the trapezoidal mean
// over the phases) and the height hh = eta(crest) - eta(trough)
void sf_psieta(Stream &s)
{
    const double rl = s.a[s.nn1 - 1];
    s.rk = s.pi2 / rl;
    s.b = rl / s.t - (s.u + s.pseta / s.d);
    s.pseta = s.a[s.nn2 - 1];
    s.eta = s.h / 2.0;
    double sum = 0.0, etac = 0.0, etat = 0.0;
    for (int i = 1; i <= s.imax1; i++) {
        s.sit = s.si[i - 1];
        s.cot = s.co[i - 1];
        sf_caleta(s);
        if (s.iecon != 0) return;
        double sitn = s.sit, cotn = s.cot;
        const double rkdeta = s.rk * (s.d + s.eta);
        const double sh1 = std::sinh(rkdeta), ch1 = std::cosh(rkdeta);
        double shn = sh1, chn = ch1;
        double shc = 0.0;
        for (int n = 1; n <= s.nn; n++) {
            shc = shc + s.a[n - 1] * shn * cotn;
            double c = cotn * s.cot - sitn * s.sit;
            sitn = sitn * s.cot + cotn * s.sit;
            cotn = c;
            c = chn * ch1 + shn * sh1;
            shn = shn * ch1 + chn * sh1;
            chn = c;
        }
        if (i == 1) {
            etac = s.eta;
            shc = shc / 2.0;
        } else if (i == s.imax1) {
            etat = s.eta;
            shc = shc / 2.0;
        }
        sum = sum + shc;
    }
    s.hh = etac - etat;
    s.pseta = sum / (double)(float)s.imax;
    s.a[s.nn2 - 1] = s.pseta;
}

// vf_dife (vf_wsfmb2.f:514-664): the error of the surface conditions, its
// gradient ekl and Gauss-Newton matrix ek over the coefficients, and the
// derivative of eta (detaa) for the next pass
void sf_dife(Stream &s)
{
    const int nn = s.nn, nn1 = s.nn1;
    double sdetaa[23], etaa[23], psza[23], psxa[23], shcn[23], qqk[23];
    const double rl = s.a[nn1 - 1];
    s.rk = s.pi2 / rl;
    s.b = rl / s.t - (s.u + s.pseta / s.d);
    s.pseta = s.a[s.nn2 - 1];
    for (int k = 1; k <= nn1; k++) {
        sdetaa[k - 1] = 0.0;
        s.ekl[k - 1] = 0.0;
        for (int l = 1; l <= nn1; l++) s.EK(k, l) = 0.0;
    }
    double se1 = 0.0;
    s.e1 = (double)2.0e10f;
    for (int i = 1; i <= s.imax1; i++) {
        s.sit = s.si[i - 1];
        s.cot = s.co[i - 1];
        sf_caleta(s);
        if (s.iecon != 0) return;
        double sitn = s.sit, cotn = s.cot;
        const double deta = s.d + s.eta;
        const double rkdeta = s.rk * deta;
        const double sh1 = std::sinh(rkdeta), ch1 = std::cosh(rkdeta);
        double shn = sh1, chn = ch1;
        double chc1 = 0.0, shs1 = 0.0, shc2 = 0.0, chs2 = 0.0, hchc1 = 0.0, hshc2 = 0.0, hchs2 = 0.0;
        for (int n = 1; n <= nn; n++) {
            const double rkn = (double)(float)n * s.rk;
            psza[n - 1] = rkn * chn * cotn;
            psxa[n - 1] = -(rkn * shn * sitn);
            shcn[n - 1] = shn * cotn;
            const double rkna = rkn * s.a[n - 1];
            const double ach = rkna * chn, ash = rkna * shn;
            const double achc = ach * cotn, ashs = ash * sitn;
            const double ashc = rkn * ash * cotn, achs = rkn * ach * sitn;
            chc1 = chc1 + achc;
            shs1 = shs1 + ashs;
            shc2 = shc2 + ashc;
            chs2 = chs2 + achs;
            hchc1 = hchc1 + achc * deta;
            hshc2 = hshc2 + ashc * deta;
            hchs2 = hchs2 + achs * deta;
            double c = cotn * s.cot - sitn * s.sit;
            sitn = sitn * s.cot + cotn * s.sit;
            cotn = c;
            c = chn * ch1 + shn * sh1;
            shn = shn * ch1 + chn * sh1;
            chn = c;
        }
        const double psz = chc1 + s.b, psx = -shs1, psze = shc2, psxe = -chs2;
        psza[nn1 - 1] = 1.0 / s.t - (chc1 + hshc2) / rl;
        psxa[nn1 - 1] = (shs1 + hchs2) / rl;
        shcn[nn1 - 1] = s.eta / s.t - hchc1 / rl;
        const double qq = (psz * psz + psx * psx) / 2.0 + s.g * s.eta - s.bq;
        const double qe = psz * psze + psx * psxe + s.g;
        const double fac = (i == 1 || i == s.imax1) ? 0.5 : 1.0;
        for (int k = 1; k <= nn1; k++) {
            etaa[k - 1] = (s.detaa[k - 1] - shcn[k - 1]) / psz;
            const double qqkk = psz * psza[k - 1] + psx * psxa[k - 1] + qe * etaa[k - 1];
            qqk[k - 1] = qqkk;
            s.ekl[k - 1] = s.ekl[k - 1] + fac * 2.0 * qq * qqkk;
            for (int l = 1; l <= k; l++) s.EK(k, l) = s.EK(k, l) + fac * 2.0 * qqkk * qqk[l - 1];
        }
        se1 = se1 + fac * qq * qq;
        for (int n = 1; n <= nn; n++) sdetaa[n - 1] = sdetaa[n - 1] + fac * (chc1 * etaa[n - 1] + shcn[n - 1]);
        sdetaa[nn1 - 1] = sdetaa[nn1 - 1] + fac * (chc1 * etaa[nn1 - 1] - hchc1 / rl);
    }
    const double fim = (double)(float)s.imax;
    for (int k = 1; k <= nn1; k++) {
        s.detaa[k - 1] = sdetaa[k - 1] / fim;
        s.ekl[k - 1] = s.ekl[k - 1] / fim;
        for (int l = 1; l <= k; l++) {
            s.EK(k, l) = s.EK(k, l) / fim;
            s.EK(l, k) = s.EK(k, l);
        }
    }
    s.e1 = std::sqrt(se1 / fim) / (s.g * s.h);
}

// vf_ludcmp / vf_lubksb (vf_wsfmb2.f:943-1032): LU decomposition with
// implicit partial pivoting (the last row of the largest scaled pivot) and
// the substitutions, on the leading n x n of a 23 x 23 column-major matrix
void sf_ludcmp(Stream &s, int n, int *indx, double tiny)
{
    double vv[23];
    for (int i = 1; i <= n; i++) {
        double aamax = 0.0;
        for (int j = 1; j <= n; j++)
            if (std::fabs(s.EK(i, j)) > aamax) aamax = std::fabs(s.EK(i, j));
        vv[i - 1] = 1.0 / aamax;
    }
    for (int j = 1; j <= n; j++) {
        for (int i = 1; i <= j - 1; i++) {
            double sum = s.EK(i, j);
            for (int k = 1; k <= i - 1; k++) sum = sum - s.EK(i, k) * s.EK(k, j);
            s.EK(i, j) = sum;
        }
        double aamax = 0.0;
        int imax = j;
        for (int i = j; i <= n; i++) {
            double sum = s.EK(i, j);
            for (int k = 1; k <= j - 1; k++) sum = sum - s.EK(i, k) * s.EK(k, j);
            s.EK(i, j) = sum;
            const double dum = vv[i - 1] * std::fabs(sum);
            if (dum >= aamax) {
                imax = i;
                aamax = dum;
            }
        }
        if (j != imax) {
            for (int k = 1; k <= n; k++) {
                const double dum = s.EK(imax, k);
                s.EK(imax, k) = s.EK(j, k);
                s.EK(j, k) = dum;
            }
            vv[imax - 1] = vv[j - 1];
        }
        indx[j - 1] = imax;
        if (s.EK(j, j) == 0.0) s.EK(j, j) = tiny;
        if (j != n) {
            const double dum = 1.0 / s.EK(j, j);
            for (int i = j + 1; i <= n; i++) s.EK(i, j) = s.EK(i, j) * dum;
        }
    }
}

void sf_lubksb(Stream &s, int n, const int *indx, double *b)
{
    int ii = 0;
    for (int i = 1; i <= n; i++) {
        const int ll = indx[i - 1];
        double sum = b[ll - 1];
        b[ll - 1] = b[i - 1];
        if (ii != 0) {
            for (int j = ii; j <= i - 1; j++) sum = sum - s.EK(i, j) * b[j - 1];
        } else if (sum != 0.0) {
            ii = i;
        }
        b[i - 1] = sum;
    }
    for (int i = n; i >= 1; i--) {
        double sum = b[i - 1];
        for (int j = i + 1; j <= n; j++) sum = sum - s.EK(i, j) * b[j - 1];
        b[i - 1] = sum / s.EK(i, i);
    }
}

// vf_lstsq (vf_wsfmb2.f:364-466): Gauss-Newton on the coefficients for the
// current Bernoulli constant; iconv = 0 when converged.  A first pass that
// halves (eta lost, or the error grown) reads the reference's da before it
// is set (vf_wsfmb2.f:446-450); here it reads 0.
void sf_lstsq(Stream &s)
{
    double da[23] = {0.0};
    int nans[23];
    s.itera = 0;
    s.hh = (double)1.0e10f;
    s.iconv = 1;
    int ifac = 1;
    for (; ifac <= s.irtry; ifac++) {
        sf_psieta(s);
        if (s.iecon == 0) break;
        sf_fact(s);
    }
    if (ifac > s.irtry) {
        s.iecon = 2;
        return;
    }
    const double sda1 = (double)1.0e10f;
    double e11 = (double)1.0e10f;
    for (s.itera = 1; s.itera <= s.iteram; s.itera++) {
        s.iconv = 1;
        bool halve = s.iecon != 0;
        if (!halve) {
            sf_dife(s);
            halve = s.e1 > e11;
        }
        if (!halve) {
            const double eps = s.a[s.nn1 - 1] * 1.0e-14;
            sf_ludcmp(s, s.nn1, nans, eps);
            sf_lubksb(s, s.nn1, nans, s.ekl);
            s.iconv = 0;
            double sda = 0.0;
            for (int k = 1; k <= s.nn1; k++) {
                const double f = std::fabs(s.ekl[k - 1] / s.depsa[k - 1]);
                sda = sda + f;
                if (f >= 1.0) s.iconv = 1;
            }
            double facd = 1.0;
            if (sda > sda1 * 0.5 && s.itera > 4) facd = 0.5;
            for (int k = 1; k <= s.nn1; k++) {
                da[k - 1] = facd * s.ekl[k - 1];
                s.a[k - 1] = s.a[k - 1] - da[k - 1];
            }
            e11 = s.e1;
        } else {
            // a step that made the error larger (or lost eta) is halved back
            s.iconv = 0;
            for (int k = 1; k <= s.nn1; k++) {
                da[k - 1] = da[k - 1] / 2.0;
                s.a[k - 1] = s.a[k - 1] + da[k - 1];
                if (std::fabs(da[k - 1]) >= s.depsa[k - 1]) s.iconv = 1;
            }
        }
        const double pseta0 = s.a[s.nn2 - 1];
        sf_psieta(s);
        if (s.iecon != 0) continue;
        if (std::fabs(s.a[s.nn2 - 1] - pseta0) >= s.depsa[s.nn2 - 1]) s.iconv = 1;
        if (s.iconv == 0) return;
    }
    s.iconv = 1;
    s.itera = s.iteram + 1;
}

// VF_SFMB02C (vf_wsfmb2.f:24-222): the wave of order n1 in depth d1, period
// t1, height h1 on a current u1 (m, s).  Returns ICONV1: 0 converged, -1
// input out of range, otherwise iecon*10000 + iterations*100 + Gauss-Newton
// passes (not converged); *rl1 is the wave length (single precision).  A
// first outer pass that produces a negative height reads the reference's
// ifg0 before it is set (vf_wsfmb2.f:123); here it reads 0.
int sf_setup(Stream &s, int n1, float d1, float t1, float h1, float u1, float *rl1)
{
    sf_prep(s);
    s.pseta = 0.0;
    s.xl = 1.0e10f;
    s.xl2 = 2.0e10f;
    s.z = 1.0e10f;
    s.ax1 = 0.0;
    s.uu1 = s.ww1 = s.uut1 = s.wwt1 = s.pp1 = 0.0;
    s.nn = n1;
    s.d = d1;
    s.t = t1;
    s.h = h1;
    s.u = u1;
    s.hh = 0.0;
    s.iecon = 0;
    if (s.nn < 1 || s.nn > 22) s.nn = 5;
    *rl1 = 0.0f;
    if (s.d <= 0.0 || s.t <= 0.0 || s.h <= 0.0) return -1;
    const double c0 = s.g * s.t / s.pi2;
    const double rlo = c0 * s.t;
    if (s.h / rlo > (double)0.3f) return -1;
    if (s.h / s.d > 1.5) return -1;
    if (s.u < -c0 / 2.0) return -1;
    if (s.d > 2.0 * rlo) s.d = 2.0 * rlo;
    sf_init(s);
    s.iconv = 1;
    s.bq = s.q0 * (double)1.01f;
    double dq = s.deltq0, rdiv = 1.0;
    sf_inita(s);
    for (int k = 1; k <= s.nn2; k++) s.a2[k - 1] = s.a[k - 1];
    int ifg0 = 0, ifg1 = 0, iterh = 1;
    double hhh0 = 0.0;
    bool conv = false;
    for (; iterh <= s.iterhm; iterh++) {
        sf_lstsq(s);
        if (iterh == 1) hhh0 = s.hh - s.h;
        const double hhh1 = s.hh - s.h;
        // the flow of labels 74, 71, 72 (height too small: raise Q) and
        // 21, 81, 82 (too large, or not converged: lower Q and restore A)
        int path;   // 74, 21
        if (s.iconv != 0) {
            path = 21;
        } else if (std::fabs(s.hh - s.h) < s.depsh) {
            conv = true;
            break;
        } else {
            path = hhh1 > 0.0 ? 21 : 74;
        }
        if (path == 21 && s.iconv != 0 && hhh1 < 0.0) path = 74;
        if (path == 74) {
            const bool to71 = iterh == 1 || hhh0 > 0.0 || hhh0 < hhh1;
            if (!to71) {
                dq = dq / 2.0;
                if (ifg0 == 1) {
                    s.bq = s.bq - dq;
                    ifg1 = 0;
                } else {
                    s.bq = s.bq + dq;
                    ifg1 = 1;
                }
            } else if (s.hh < 0.0) {
                if (ifg0 == 1) {
                    dq = dq / 2.0;
                    s.bq = s.bq - dq;
                    ifg1 = 0;
                    if (iterh == 1) ifg0 = 0;
                } else {
                    s.bq = s.bq + dq;
                    ifg1 = 1;
                    if (iterh == 1) ifg0 = 1;
                }
            } else {
                s.bq = s.bq + dq;
                ifg1 = 1;
                if (iterh == 1) ifg0 = 1;
            }
            for (int k = 1; k <= s.nn2; k++) s.a2[k - 1] = s.a[k - 1];
        } else {
            const bool to81 = iterh == 1 || hhh0 < 0.0 || hhh0 > hhh1;
            if (!to81) {
                dq = dq / 2.0;
                if (ifg0 == 0) {
                    s.bq = s.bq + dq;
                    ifg1 = 1;
                } else {
                    s.bq = s.bq - dq;
                    ifg1 = 0;
                }
            } else {
                rdiv = 2.0;
                dq = dq / rdiv;
                s.bq = s.bq - dq;
                ifg1 = 0;
                if (iterh == 1) ifg0 = 0;
            }
            for (int k = 1; k <= s.nn2; k++) s.a[k - 1] = s.a2[k - 1];
        }
        if (iterh == 1) continue;
        hhh0 = hhh1;
        ifg0 = ifg1;
        if (dq < s.depsq) break;
    }
    (void)rdiv;
    *rl1 = (float)s.a[s.nn1 - 1];
    if (conv) return 0;
    return s.iecon * 10000 + iterh * 100 + s.itera;
}

// VF_SFMB1 (vf_wsfmb2.f:803-829): the surface level (m) at phase xl1
float sf_level(Stream &s, float xl1)
{
    const double ax = std::fabs(xl1 - s.xl);
    if (!(ax < (double)1.0e-7f)) {
        s.xl = xl1;
        const double rkx = s.pi2 * (double)s.xl;
        s.sit = std::sin(rkx);
        s.cot = std::cos(rkx);
        s.ax1 = std::fabs(s.xl2 - s.xl);
    }
    if (!(s.ax1 <= (double)1.0e-7f)) {
        s.eta = 0.0;
        sf_caleta(s);
        s.xl2 = s.xl;
    }
    return (float)s.eta;
}

// VF_SFMB2 (vf_wsfmb2.f:831-909): the horizontal and vertical velocity (m/s)
// at phase xl1 and height z1 (m from the still water level)
void sf_velocity(Stream &s, float xl1, float z1, float *u1, float *w1)
{
    bool same = false;
    bool newz = true;
    if (!(std::fabs(xl1 - s.xl) >= 1.0e-7f)) {
        if (!(std::fabs(z1 - s.z) >= 1.0e-7f)) same = true;
    } else {
        s.xl = xl1;
        const double rkx = s.pi2 * (double)s.xl;
        s.sit = std::sin(rkx);
        s.cot = std::cos(rkx);
        if (std::fabs(z1 - s.z) < 1.0e-7f) newz = false;
    }
    if (!same) {
        if (newz) {
            s.z = z1;
            double rkdz = s.rk * (s.d + (double)s.z);
            if (rkdz < 0.0) rkdz = 0.0;
            s.sh = std::sinh(rkdz);
            s.ch = std::cosh(rkdz);
        }
        double chc = 0.0, shs = 0.0, chs = 0.0, shc = 0.0;
        double sitn = s.sit, cotn = s.cot, shn = s.sh, chn = s.ch;
        for (int n = 1; n <= s.nn; n++) {
            const double an = (double)(float)n * s.rk * s.a[n - 1];
            const double ach = an * chn, ash = an * shn;
            const double sg = (double)(float)n * s.sigma;
            chc = chc + ach * cotn;
            shs = shs + ash * sitn;
            chs = chs + ach * sitn * sg;
            shc = shc + ash * cotn * sg;
            double c = cotn * s.cot - sitn * s.sit;
            sitn = sitn * s.cot + cotn * s.sit;
            cotn = c;
            c = chn * s.ch + shn * s.sh;
            shn = shn * s.ch + chn * s.sh;
            chn = c;
        }
        s.uu1 = -chc + s.u + s.pseta / s.d;
        s.ww1 = -shs;
        const double cc = s.sigma / s.rk;
        s.uut1 = -(chs * (1.0 - s.uu1 / cc)) - shc * s.ww1 / cc;
        s.wwt1 = shc * (1.0 - s.uu1 / cc) - chs * s.ww1 / cc;
        const double du = s.uu1 - cc;
        s.pp1 = (s.bq - (du * du + s.ww1 * s.ww1) / 2.0) / s.g - (double)s.z;
    }
    *u1 = (float)s.uu1;
    *w1 = (float)s.ww1;
}

// ---- VF_CWMAK0/1/2 and VF_CWZERO -----------------------------------------

struct Theory {
    int kind;   // -2 Stokes, -1 cnoidal, 1..21 the stream function of that order
    Stokes stk;
    Cnoidal cnd;
    Stream sf;
    float sf_rl;   // VF_SFMB02C's wave length (m)
};

// VF_CWMAK0 (vf_cwmak0.f): the wave of depth d, period t, height h (m, s);
// CAD_ERR_ARG when the stream function does not converge (VF_CSETUP's
// "(VF_SFMB02C) NOT CONVERGED.", vf_cwmak0.f:53-55)
int cwmak0(cad_ctx *c, Theory &w, int kind, double d, double t, double h)
{
    w.kind = kind;
    if (kind >= 1) {
        // SI units, no current (vf_cwmak0.f:43-55)
        const int icn = sf_setup(w.sf, kind, (float)d, (float)t, (float)h, 0.0f, &w.sf_rl);
        if (icn != 0) {
            cad_set_error("VF_CSETUP: (VF_SFMB02C) NOT CONVERGED. (stream function of order %d, "
                          "depth %g, period %g, height %g: code %d)", kind, d, t, h, icn);
            return CAD_ERR_ARG;
        }
        return CAD_OK;
    }
    const float r4d = (float)(d * 100.0), r4t = (float)t, r4h = (float)(h * 100.0);
    if (kind == -2)
        stk0(w.stk, c->wv_stk_dddk, 5, r4d, r4t, r4h);
    else
        cnd0(w.cnd, 3, r4d, r4t, r4h);
    return CAD_OK;
}

// the wave length (m) VF_CWMAK0 returns (WVLENG)
double cwlength(const Theory &w)
{
    if (w.kind >= 1) return (double)w.sf_rl;
    return (double)(w.kind == -2 ? w.stk.rl : w.cnd.rl) * 0.01;
}

// VF_CWMAK1: the level (m) at phase wvt; for the stream function also the
// VF_SFMB2 call at phase -1 that follows it (vf_cwmak1.f:33-37), which
// resets VF_SFMB2's cached phase
double cwmak1(Theory &w, double wvt)
{
    const float r4tn = (float)wvt;
    if (w.kind >= 1) {
        const float z = sf_level(w.sf, r4tn);
        float u, v;
        sf_velocity(w.sf, -1.0f, 0.0f, &u, &v);
        return (double)z;
    }
    const float z = w.kind == -2 ? stk1(w.stk, r4tn) : cnd1(w.cnd, r4tn);
    return (double)z * 0.01;
}

// VF_CWMAK2: the normal and tangential velocity (m/s) at phase wvt, height zc
// (m, from the still water level)
void cwmak2(Theory &w, double wvt, double zc, double *un, double *ut)
{
    const float r4tn = (float)wvt;
    float u, v;
    if (w.kind >= 1) {
        sf_velocity(w.sf, r4tn, (float)zc, &u, &v);
        *un = (double)u;
        *ut = (double)v;
        return;
    }
    const float r4z = (float)(zc * 100.0);
    if (w.kind == -2)
        stk2(w.stk, r4tn, r4z, &u, &v);
    else
        cnd2(w.cnd, r4tn, r4z, &u, &v);
    *un = (double)u * 0.01;
    *ut = (double)v * 0.01;
}

// VF_CWZERO (vf_cwzero.f): the phase in (0.5, 1] where the level crosses zero,
// by 100 steps down from 1 then 20 bisections (|level| < 1e-8, 1e-6 for the
// stream function, ends early)
double cwzero(Theory &w)
{
    const float eps = w.kind >= 1 ? 1.0e-6f : 1.0e-8f;
    double wvzero = 0.5;
    float tn1 = 1.0f;
    float z1 = (float)cwmak1(w, (double)tn1);
    if (std::fabs(z1) < eps) return (double)tn1;
    float tn2 = tn1, z2 = z1;
    for (int l = 1; l <= 100; l++) {
        tn2 = 1.0f - 0.5f * (float)l / 100.0f;
        z2 = (float)cwmak1(w, (double)tn2);
        if (z1 * z2 <= 0.0f) break;
        tn1 = tn2;
        z1 = z2;
    }
    if (std::fabs(z2) < eps) return (double)tn2;
    if (tn1 == tn2) return wvzero;
    for (int l = 1; l <= 20; l++) {
        const float tn = (tn1 + tn2) * 0.5f;
        wvzero = (double)tn;
        const float z = (float)cwmak1(w, wvzero);
        if (std::fabs(z) < eps) break;
        if (z1 * z >= 0.0f) {
            tn1 = tn;
            z1 = z;
        } else {
            tn2 = tn;
            z2 = z;
        }
    }
    return wvzero;
}

// level scale of each column next to the side (vf_bwuwn.f:300-330)
__global__ void __launch_bounds__(64) k_wave_scale(Dims d, const double *__restrict__ zz,
                                                   const int *__restrict__ nf,
                                                   const double *__restrict__ ff, int xside, int cpos,
                                                   int q0, int nq, double dep, double wvlvl, double *vws)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const int q = q0 + t;
    double val = 0.0;
    for (int k = 2; k <= d.nk - 1; k++) {
        const long long c = xside ? cid(d, cpos, q, k) : cid(d, q, cpos, k);
        const int nw = nf[c];
        const double dz = zz[(long long)(k - 1) * 6 + 1];
        if (nw == -1) {
            val = val + dz * 1.0;
        } else if (nw == 0) {
            val = val + dz * ff[c];
        } else if (nw != 8) {
            val = val + dz * ff[c];
            break;
        }
    }
    val = val - (wvlvl - zz[6]);
    // vws holds the column's expected level on entry
    vws[t] = (vws[t] + dep) / (val + dep);
}

// the side's boundary faces from the host's velocities (vf_bwuwn.f:331-380):
// uv holds UN then UT for faces k = 1..NUMK of each column
__global__ void __launch_bounds__(64) k_wave_theory_faces(Dims d, const int *__restrict__ indf,
                                                          const double *__restrict__ vws,
                                                          const double *__restrict__ uv, int xside,
                                                          int pos, int q0, int nq, double aw,
                                                          double sa, double sb, double *vel,
                                                          double *bcu, double *bcv, double *bcw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const int q = q0 + t;
    const double s = vws[t];
    const double *col = uv + (size_t)t * 2 * d.nk;
    for (int k = 2; k <= d.nk - 1; k++) {
        const long long f = xside ? cid(d, pos, q, k) : cid(d, q, pos, k);
        const int l = indf[f];
        if (l < 1) continue;
        const double un = col[k - 1], ut = col[d.nk + k - 1];
        const double u = un * s * aw * sa;
        const double v = un * s * aw * sb;
        bcu[l - 1] = u;
        bcv[l - 1] = v;
        bcw[l - 1] = ut * s * aw;
        vel[f] = xside ? u : v;
    }
}

}  // namespace

int cad_wave_theory_eval(int func, double depth, double height, double period, float *stk_state,
                         double wvt, int nz, const double *zc, double *un, double *ut,
                         double *level, double *t0, double *length)
{
    if ((func != -2 && func != -1 && (func < 1 || func > 21)) || !stk_state || (nz > 0 && (!zc || !un || !ut)) ||
        nz < 0) {
        cad_set_error("cad_wave_theory_eval: func -2 (Stokes), -1 (cnoidal) or 1..21 (stream function "
                      "order), the state, and nz >= 0 heights with their results");
        return CAD_ERR_ARG;
    }
    if (!(depth >= ZERO_REF) || !(height >= ZERO_REF) || !(period >= ZERO_REF)) {
        cad_set_error("cad_wave_theory_eval: depth, height and period must be positive");
        return CAD_ERR_ARG;
    }
    Theory w;
    w.kind = func;
    const float r4d = (float)(depth * 100.0), r4t = (float)period, r4h = (float)(height * 100.0);
    if (func >= 1) {
        const int rc = cwmak0(nullptr, w, func, depth, period, height);
        if (rc) return rc;
    } else if (func == -2) {
        stk0(w.stk, *stk_state, 5, r4d, r4t, r4h);
    } else {
        cnd0(w.cnd, 3, r4d, r4t, r4h);
    }
    if (length) *length = cwlength(w);
    if (t0) *t0 = cwzero(w);
    if (level) *level = cwmak1(w, wvt);
    for (int k = 0; k < nz; k++) cwmak2(w, wvt, zc[k], &un[k], &ut[k]);
    return CAD_OK;
}

void cad_wave_theory_free(cad_ctx *c)
{
    if (c->wv_vws) hipFree(c->wv_vws);
    if (c->wv_uv) hipFree(c->wv_uv);
    if (c->wv_vws_host) hipHostFree(c->wv_vws_host);
    if (c->wv_uv_host) hipHostFree(c->wv_uv_host);
    std::free(c->wv_ph_host);
    c->wv_vws = c->wv_uv = c->wv_vws_host = c->wv_uv_host = c->wv_ph_host = nullptr;
    c->wv_cols = 0;
}

int cad_wave_theory_side(cad_ctx *c, int side, int func, int l1, int l2, double depth, double height,
                         double period, double ampl, double wvlvl, double angle, double ax, double ay)
{
    if (side < 1 || side > 4) {
        cad_set_error("cad_set_wave_theory: side %d (1..4 = X-, X+, Y-, Y+)", side);
        return CAD_ERR_ARG;
    }
    if (func < -2 || func > 21) {
        cad_set_error("cad_set_wave_theory: function %d (0 STK-CND, -1 CNOIDAL, -2 STOKES, 1..21 STREAM "
                      "of that order; the reference's order 22 writes past its own arrays)", func);
        return CAD_ERR_ARG;
    }
    if (c->op_on[side - 1]) {
        cad_set_error("cad_set_wave_theory: side %d is a radiation side (DEFINED DOUBLY.)", side);
        return CAD_ERR_ARG;
    }
    // VF_CSETUP (vf_csetup.f:86-94): depth, height and period are needed
    if (!(depth >= ZERO_REF) || !(height >= ZERO_REF) || !(period >= ZERO_REF)) {
        cad_set_error("cad_set_wave_theory: (WAVE-BC DEPTH / HEIGHT / PERIOD) NOT FOUND.");
        return CAD_ERR_ARG;
    }
    const int lim = side <= 2 ? c->d.nj : c->d.ni;
    if (l1 < 2 || l2 > lim - 1 || l1 > l2) {
        cad_set_error("cad_set_wave_theory: columns %d..%d outside 2..%d", l1, l2, lim - 1);
        return CAD_ERR_ARG;
    }
    if (!std::isfinite(angle) || !std::isfinite(ax) || !std::isfinite(ay)) {
        cad_set_error("cad_set_wave_theory: the angle and its origin must be finite");
        return CAD_ERR_ARG;
    }
    // STK-CND: Stokes up to an Ursell number of 25 (vf_csetup.f:101-111)
    int kind = func;
    if (kind == 0) {
        const double wus = c->prm.grz0 * height * period * period / depth / depth;
        kind = wus <= 25.0 ? -2 : -1;
    }
    CAD_CHECK(hipSetDevice(c->device));
    if (!c->wv_vws) {
        const size_t cols = (size_t)(c->d.ni > c->d.nj ? c->d.ni : c->d.nj);
        const size_t nuv = 4 * cols * 2 * (size_t)c->d.nk;
        CAD_CHECK(hipMalloc((void **)&c->wv_vws, 4 * cols * sizeof(double)));
        CAD_CHECK(hipMalloc((void **)&c->wv_uv, nuv * sizeof(double)));
        CAD_CHECK(hipHostMalloc((void **)&c->wv_vws_host, 4 * cols * sizeof(double), hipHostMallocDefault));
        CAD_CHECK(hipHostMalloc((void **)&c->wv_uv_host, nuv * sizeof(double), hipHostMallocDefault));
        c->wv_ph_host = (double *)std::malloc(4 * cols * sizeof(double));
        if (!c->wv_ph_host) {
            cad_set_error("cad_set_wave_theory: out of host memory");
            return CAD_ERR_HIP;
        }
        c->wv_cols = cols;
    }
    if (!c->wv_err) {
        CAD_CHECK(hipMalloc((void **)&c->wv_err, 64));
        CAD_CHECK(hipHostMalloc((void **)&c->wv_err_host, 64, hipHostMallocDefault));
        CAD_CHECK(hipMemset(c->wv_err, 0, 64));
        c->wv_err_host[0] = 0;
    }
    // VF_CSETUP: the wave once (VF_STK0 carries its secant step on), its
    // length, and the phase origin where the level crosses zero
    // (vf_csetup.f:113-125)
    Theory w;
    RET_IF_W(cwmak0(c, w, kind, depth, period, height));
    const int q = side - 1;
    c->wv_on[q] = 1;
    c->wv_kind[q] = kind;
    c->wv_l1[q] = l1;
    c->wv_l2[q] = l2;
    c->wv_d[q] = depth;
    c->wv_h[q] = height;
    c->wv_t[q] = period;
    c->wv_a[q] = ampl;
    c->wv_wl[q] = cwlength(w);
    c->wv_ang[q][0] = angle;
    c->wv_ang[q][1] = ax;
    c->wv_ang[q][2] = ay;
    c->wv_t0s[q] = cwzero(w);
    c->wv_wvlvl = wvlvl;
    c->wv_level[q] = 0.0;
    return CAD_OK;
}

int cad_wave_theory_apply(cad_ctx *c, int jd)
{
    const Dims &d = c->d;
    const int q = jd - 1;
    const double dep = c->wv_d[q], t = c->wv_t[q], t0 = c->wv_t0s[q];
    double a = c->wv_a[q];
    // phase in single precision and ramp (vf_bwuwn.f:270-283)
    float r4tn = (float)(t0 - c->prm.tnow / t);
    r4tn = r4tn - (float)(int)r4tn;
    if (r4tn < 0.0f) r4tn = r4tn + 1.0f;
    double aw = 1.0;
    if (a >= ZERO_REF) {
        a = c->prm.tnow / t / a;
        if (a < 1.0) aw = 0.5 * std::sin(PI_REF * (a - 0.5)) + 0.5;
    }
    Theory w;
    RET_IF_W(cwmak0(c, w, c->wv_kind[q], dep, t, c->wv_h[q]));
    const double wvt = (double)r4tn;
    const double wvz = cwmak1(w, wvt);
    c->wv_level[q] = wvz * aw;
    // direction of the wave (vf_bwuwn.f:297-325): normal to the side, or at
    // ANGLE degrees from it with the phase shifted along the ray through
    // (SX, SY)
    double sa, sb, sc = 0.0;
    const int isa = std::fabs(c->wv_ang[q][0]) > ZERO_REF;
    if (!isa) {
        sa = jd == 1 ? 1.0 : (jd == 2 ? -1.0 : 0.0);
        sb = jd == 3 ? 1.0 : (jd == 4 ? -1.0 : 0.0);
    } else {
        double ang = c->wv_ang[q][0];
        if (jd == 2) ang = ang + 180.0;
        else if (jd == 3) ang = ang + 90.0;
        else if (jd == 4) ang = ang + 270.0;
        ang = ang * PI_REF / 180.0;
        sb = std::sin(ang);
        sa = std::cos(ang);
        sc = -(sa * c->wv_ang[q][1] + sb * c->wv_ang[q][2]);
    }
    const int xside = jd <= 2;
    const int pos = xside ? (jd == 1 ? 2 : d.ni) : (jd == 3 ? 2 : d.nj);
    const int cpos = xside ? (jd == 1 ? 2 : d.ni - 1) : (jd == 3 ? 2 : d.nj - 1);
    const int q0 = c->wv_l1[q], nq = c->wv_l2[q] - c->wv_l1[q] + 1;
    if (nq <= 0 || (size_t)nq > c->wv_cols) {
        cad_set_error("cad_wave_theory_apply: side %d has %d columns (buffers hold %zu)", jd, nq,
                      c->wv_cols);
        return CAD_ERR_STATE;
    }
    const double wvlvl = c->wv_wvlvl;
    double *vws = c->wv_vws + (size_t)q * c->wv_cols;
    double *vws_h = c->wv_vws_host + (size_t)q * c->wv_cols;
    const size_t per = 2 * (size_t)d.nk;
    double *uv = c->wv_uv + (size_t)q * c->wv_cols * per;
    double *uv_h = c->wv_uv_host + (size_t)q * c->wv_cols * per;
    // each column's phase and expected level; the scale kernel turns the
    // level into the column's VWS in place
    double *ph = c->wv_ph_host + (size_t)q * c->wv_cols;
    const double *xh = c->xx_host, *yh = c->yy_host;
    for (int tq = 0; tq < nq; tq++) {
        double wq = wvt, zq = wvz;
        if (isa) {
            const int qq = q0 + tq;
            double x0, y0;
            if (xside) {
                x0 = xh[(size_t)(pos - 1) * 6];
                y0 = (yh[(size_t)(qq - 1) * 6] + yh[(size_t)qq * 6]) * 0.5;
            } else {
                y0 = yh[(size_t)(pos - 1) * 6];
                x0 = (xh[(size_t)(qq - 1) * 6] + xh[(size_t)qq * 6]) * 0.5;
            }
            const double rr = (sa * x0 + sb * y0 + sc) / std::sqrt(sa * sa + sb * sb);
            float r4 = (float)(t0 - (c->prm.tnow / t - rr / c->wv_wl[q]));
            r4 = r4 - (float)(int)r4;
            if (r4 < 0.0f) r4 = r4 + 1.0f;
            wq = (double)r4;
            zq = cwmak1(w, wq);
        }
        ph[tq] = wq;
        vws_h[tq] = zq;
    }
    CAD_CHECK(hipMemcpyAsync(vws, vws_h, nq * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_wave_scale, dim3((nq + 63) / 64), dim3(64), 0, c->stream, d, c->zz, c->f.nf,
                       c->f.ff, xside, cpos, q0, nq, dep, wvlvl, vws);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(vws_h, vws, nq * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    // VF_CWMAK2 at each face height of each column (ZZ(1,K): the host's
    // copy of the z metric)
    const double *z = c->zz_host;
    for (int tq = 0; tq < nq; tq++) {
        const double s = vws_h[tq];
        double *col = uv_h + (size_t)tq * per;
        for (int k = 2; k <= d.nk - 1; k++) {
            double zc = (z[(size_t)(k - 1) * 6] + z[(size_t)k * 6]) * 0.5 - wvlvl;
            zc = s * (zc + dep) - dep;
            cwmak2(w, ph[tq], zc, &col[k - 1], &col[d.nk + k - 1]);
        }
    }
    CAD_CHECK(hipMemcpyAsync(uv, uv_h, (size_t)nq * per * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    hipLaunchKernelGGL(k_wave_theory_faces, dim3((nq + 63) / 64), dim3(64), 0, c->stream, d,
                       xside ? c->indx : c->indy, vws, uv, xside, pos, q0, nq, aw, sa, sb,
                       xside ? c->f.uu : c->f.vv, c->f.bcu, c->f.bcv, c->f.bcw);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}
