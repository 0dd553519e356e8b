// cad_wave_theory.hip -- the wave-maker's theory functions: MODEL WAVE-BC
// <side> FUNC STOKES (fifth order), CNOIDAL (third order) and STK-CND (one of
// the two by the Ursell number), applied in VF_BWUWN (vf_bwuwn.f:257-480)
// through VF_CWMAK0/1/2 (vf_cwmak0.f, vf_cwmak1.f, vf_cwmak2.f).
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

// ---- VF_CWMAK0/1/2 and VF_CWZERO -----------------------------------------

struct Theory {
    int kind;   // -2 Stokes, -1 cnoidal
    Stokes stk;
    Cnoidal cnd;
};

// VF_CWMAK0 (vf_cwmak0.f): the wave of depth d, period t, height h (m, s)
void cwmak0(cad_ctx *c, Theory &w, int kind, double d, double t, double h)
{
    w.kind = kind;
    const float r4d = (float)(d * 100.0), r4t = (float)t, r4h = (float)(h * 100.0);
    if (kind == -2)
        stk0(w.stk, c->wv_stk_dddk, 5, r4d, r4t, r4h);
    else
        cnd0(w.cnd, 3, r4d, r4t, r4h);
}

// the wave length (m) VF_CWMAK0 returns (WVLENG)
double cwlength(const Theory &w)
{
    return (double)(w.kind == -2 ? w.stk.rl : w.cnd.rl) * 0.01;
}

// VF_CWMAK1: the level (m) at phase wvt
double cwmak1(const Theory &w, double wvt)
{
    const float r4tn = (float)wvt;
    const float z = w.kind == -2 ? stk1(w.stk, r4tn) : cnd1(w.cnd, r4tn);
    return (double)z * 0.01;
}

// VF_CWMAK2: the normal and tangential velocity (m/s) at phase wvt, height zc
// (m, from the still water level)
void cwmak2(const Theory &w, double wvt, double zc, double *un, double *ut)
{
    const float r4tn = (float)wvt;
    const float r4z = (float)(zc * 100.0);
    float u, v;
    if (w.kind == -2)
        stk2(w.stk, r4tn, r4z, &u, &v);
    else
        cnd2(w.cnd, r4tn, r4z, &u, &v);
    *un = (double)u * 0.01;
    *ut = (double)v * 0.01;
}

// VF_CWZERO (vf_cwzero.f): the phase in (0.5, 1] where the level crosses zero,
// by 100 steps down from 1 then 20 bisections (|level| < 1e-8 ends early)
double cwzero(const Theory &w)
{
    const float eps = 1.0e-8f;
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
    if ((func != -2 && func != -1) || !stk_state || (nz > 0 && (!zc || !un || !ut)) || nz < 0) {
        cad_set_error("cad_wave_theory_eval: func -2 (Stokes) or -1 (cnoidal), the state, and "
                      "nz >= 0 heights with their results");
        return CAD_ERR_ARG;
    }
    if (!(depth >= ZERO_REF) || !(height >= ZERO_REF) || !(period >= ZERO_REF)) {
        cad_set_error("cad_wave_theory_eval: depth, height and period must be positive");
        return CAD_ERR_ARG;
    }
    Theory w;
    w.kind = func;
    const float r4d = (float)(depth * 100.0), r4t = (float)period, r4h = (float)(height * 100.0);
    if (func == -2)
        stk0(w.stk, *stk_state, 5, r4d, r4t, r4h);
    else
        cnd0(w.cnd, 3, r4d, r4t, r4h);
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
    if (func < -2 || func > 0) {
        cad_set_error("cad_set_wave_theory: function %d (0 STK-CND, -1 CNOIDAL, -2 STOKES)", func);
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
    cwmak0(c, w, kind, depth, period, height);
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
    cwmak0(c, w, c->wv_kind[q], dep, t, c->wv_h[q]);
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
