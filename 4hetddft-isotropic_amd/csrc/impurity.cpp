// impurity.cpp -- impurity-helium pair potentials of Select_Pot
// (V_impur.f90:84-929), evaluated on the host once per run to fill the
// impurity table (potenimpini.f90:11-25).
//
// The evaluation follows the reference's arithmetic so that the table agrees
// with it to the last bits:
//  - real**integer is binary powering (ipow) for a variable exponent or a
//    constant above 8, and a chain of multiplications (cpow) for a constant
//    exponent up to 8, which is what the Fortran compiler emits;
//  - constants the reference writes as default-real literals (no d0) are
//    single-precision values widened to double: they appear here as
//    (double)<literal>f;
//  - the inverse-power fits declared Real*16 are summed in binary128
//    (__float128), the rest in binary64.
// Built by the host C++ compiler (-std=gnu++17 for the Q literal suffix).
#include "impurity.hpp"

#include <cmath>
#include <cstring>
#include <string>

namespace hetd {
namespace imp {
namespace {

typedef __float128 quad;

template <typename T> T ipow(T x, int n) {
    unsigned m = n < 0 ? 0u - (unsigned)n : (unsigned)n;
    T y = (m & 1u) ? x : T(1);
    while (m >>= 1) {
        x = x * x;
        if (m & 1u) y = y * x;
    }
    return n < 0 ? T(1) / y : y;
}

// x**N for a constant N: the Fortran compiler multiplies N <= 8 out as
// x*(x*(...*x)) and uses binary powering above
template <int N, typename T> T cpow(T x) {
    if constexpr (N <= 8) {
        T y = x;
        for (int i = 1; i < N; ++i) y = x * y;
        return y;
    } else {
        return ipow(x, N);
    }
}

inline double sq(double x) { return x * x; }

// ------------------------------------------------ inverse-power fits
// V = fcutoff (x <= rcutoff); |aa x^2 + bb x + cc| (x <= r0); else
//   A: sum Pg(k) xq^(k+k0),            xq = 1/x
//   B: sum Pg(k) xq^(k0+k)  with the power carried from term to term
//   C: sum Pg(k) / x^(k+5)  (x^(k+5) in binary64)
// (V_impur.f90, e.g. V_LJ_OT :1009, V_Ba_plus_2D0 :1235, V_Rb_6s :1898)
struct InvPowFit {
    const char* selec;
    char style;
    bool quad_sum;
    int npg, k0;
    double rcutoff, fcutoff, r0, aa, bb, cc;
    quad pg[20];
};

const InvPowFit kFits[] = {
#include "impurity_fits.inc"
};

double inv_pow_fit(const InvPowFit& f, double x) {
    if (x <= f.rcutoff) return f.fcutoff;
    if (x <= f.r0) return std::fabs(f.aa * sq(x) + f.bb * x + f.cc);
    if (!f.quad_sum) {  // binary64 sum (V_LJ_OT): xq rounded from the binary128 quotient
        const double xq = (double)((quad)1 / (quad)x);
        double s = 0.0;
        for (int k = 1; k <= f.npg; ++k) s = s + (double)f.pg[k - 1] * ipow(xq, k + f.k0);
        return s;
    }
    quad s = 0;
    if (f.style == 'C') {
        for (int k = 1; k <= f.npg; ++k) s = s + f.pg[k - 1] / (quad)ipow(x, k + 5);
        return (double)s;
    }
    const quad xq = (quad)1 / (quad)x;
    if (f.style == 'A') {
        for (int k = 1; k <= f.npg; ++k) s = s + f.pg[k - 1] * ipow(xq, k + f.k0);
    } else {
        quad a = ipow(xq, f.k0 + 1);
        for (int k = 1; k <= f.npg; ++k) {
            s = s + f.pg[k - 1] * a;
            a = a * xq;
        }
    }
    return (double)s;
}

// --------------------------------------------------- closed forms

// Tang-Toennies rare gas-He (V_Xe_He :2728, V_Ar_He :2678), atomic units in
// the fit, K out
double tang_toennies(double AA, double BB, double C6, double C8, double C10, double r) {
    const double rbohr = 0.5292, h2k = 27.21 * 11604.0;
    double B[19], C[8], S[19];
    B[1] = BB;
    for (int i = 2; i <= 18; ++i) B[i] = B[i - 1] * BB / (double)i;
    C[1] = C6;
    C[2] = C8;
    C[3] = C10;
    for (int i = 4; i <= 7; ++i) C[i] = cpow<3>(C[i - 1] / C[i - 2]) * C[i - 3];
    const double rr = r / rbohr, e = std::exp(-BB * rr);
    double st = rr;
    S[1] = B[1] * st;
    for (int i = 2; i <= 18; ++i) {
        st = st * rr;
        S[i] = S[i - 1] + B[i] * st;
    }
    double aux = AA * e;
    const double s1 = 1.0 / sq(rr);
    st = cpow<3>(s1);
    for (int i = 1; i <= 7; ++i) {
        aux = aux + ((S[2 * i + 4] + 1.0) * e - 1.0) * C[i] * st;
        st = st * s1;
    }
    return aux * h2k;
}

// He-He of Aziz (V_Aziz_He :967)
double aziz_he(double rr) {
    const double eps = 10.948, A = 1.8443101e5, alpha = 10.43329537, beta = -2.27965105, D = 1.4826;
    const double C6 = 1.36745214, C8 = 0.42123807, C10 = 0.17473318, rm = 2.963;
    const double r = rr / rm;
    double ff = 1.0;
    if (r <= D) ff = std::exp(-sq(D / r - 1.0));
    if (r == 0.0) return A * eps;
    return (A * std::exp(-(alpha * r) + beta * sq(r)) -
            ff * (C6 / cpow<6>(r) + C8 / cpow<8>(r) + C10 / cpow<10>(r))) * eps;
}

// Ag-He states (V_Ag_gs :1035, V_Ag_Pi :1055, V_Ag_Sig :1075):
// A exp(-alpha r - beta r^2) - f(r) (C6/r^6 + C8/r^8 + C10/r^10 + C12/r^12)
struct AgFit { double A, alpha, beta, D, C6, C8, C10, C12; };
double ag_fit(const AgFit& p, double r) {
    double ff = 1.0;
    if (r <= p.D) ff = std::exp(-sq(p.D / r - 1.0));
    if (r == 0.0) return p.A;
    return p.A * std::exp(-(p.alpha * r) - p.beta * sq(r)) -
           ff * (p.C6 / cpow<6>(r) + p.C8 / cpow<8>(r) + p.C10 / cpow<10>(r) + p.C12 / cpow<12>(r));
}
const AgFit kAgGs = {(double)78741.3877f, (double)0.65299f, (double)0.36367f, (double)12.093f,
                     137042.07, 189386.217, 111358462.0, 1.17670377e10};
const AgFit kAgPi = {0.7095092e6, (double)1.9928f, (double)0.45500f, (double)4.0562f,
                     0.21661134e6, (double)123454.7952f, 0.307155024e6, 0.740464032e-10};
const AgFit kAgSig = {(double)40005.2f, (double)0.41529f, (double)0.14888f, (double)25.730f,
                      1.97443e-9, 3.37393e-6, (double)0.00788363f, 7.52624e12};

// Patil's alkali-He (V_alka :1094): elements LI, NA, K, RB, CS
double alkali(int e, double dist) {
    const double av[5] = {1.588, 1.627, 1.771, 1.805, 1.869};
    const double bv = 0.744;
    const double c6v[5] = {22.5, 24.7, 38.9, 44.6, 51.2};
    const double c8v[5] = {1.06e3, 1.29e3, 2.66e3, 3.18e3, 4.34e3};
    const double r2 = 2.37, r4 = 7.78, r6 = 50.5;
    const double agv[5] = {0.0519, 0.0450, 0.0259, 0.0226, 0.0173};
    const double evk = (double)11604.448f, bohr = (double)0.529177249f, hartree = (double)27.211608f;
    const double factor = evk * hartree;
    const double a = av[e], b = bv, c6 = c6v[e], c8 = c8v[e], agran = agv[e];
    const double d = dist / bohr;
    const double pi = 4.0 * std::atan(1.0);
    const double abar = (a + b) / 4.0, u = a - 1.0, v = -0.5 * sq(a) * u;
    const double roa = agran * std::pow(d, 2.0 * u) * sq(1.0 + v / d) * std::exp(-(2.0 * d / a));
    double f[3];
    for (int l = 1; l <= 2; ++l) {
        double s = 0.0, fac = 1.0;
        for (int n = 0; n <= 2 * l + 7; ++n) {
            if (n >= 2) fac = fac * (double)n;
            s = s + ipow(d / abar, n) / fac;
        }
        f[l] = 1.0 - s * std::exp(-d / abar);
    }
    const double vimp = 4.0 * pi / 3.0 * roa * (r2 + 37.0 / 90.0 * r4 / sq(a) + 28.0 / 525.0 * r6 / cpow<4>(a)) -
                        c6 * f[1] / cpow<6>(d) - c8 * f[2] / cpow<8>(d);
    return vimp * factor;
}

// Ba-He (V_Ba_gs :1194)
double ba_gs(double x) {
    if (x <= 3.7) return 1297.1;
    return -6.146426949008034e14 / cpow<18>(x) + 1.6797251118695925e14 / cpow<16>(x) -
           1.6513058293510484e13 / cpow<14>(x) + 7.015334514979032e11 / cpow<12>(x) -
           1.1066880774828484e10 / cpow<10>(x) + 4.900282758445466e7 / cpow<8>(x) - 450294.6878721156 / cpow<6>(x);
}

// Ba+-He, earlier fit (Func_Ba_plus_gs :1460): even inverse powers 6..36
double ba_plus_old_gs(double x) {
    const double pg[16] = {-5.75372193971343e+06, 7.15787621605274e+08, -4.96564765588082e+10, 1.98124418444111e+12,
                           -4.71450937592845e+13, 6.52366641357938e+14, -3.61296282269969e+15, -3.34341761554747e+16,
                           6.27946406049051e+17,  -1.25469318152625e+18, -2.47420248203689e+19, -1.56040616253124e+20,
                           4.94566594588086e+21,  -2.19511834000164e+22, -6.01807193986940e+22, 4.58579988466807e+23};
    const double r0 = 3.0, aa = 1.406590885854125e+06, bb = -8.567381411545813e+06, cc = 1.304908921806432e+07;
    if (x <= r0) return aa * sq(x) + bb * x + cc;
    double s = 0.0;
    for (int k = 1; k <= 16; ++k) s = s + pg[k - 1] / ipow(x, 2 * (k - 1) + 6);
    return s;
}

// Cs(7s)-He (V_Cs_7s :1679): sum of x^(k-1) exp(-Ps(k) x) terms, x >= 2.5
double cs_7s(double xx) {
    const double pg[12] = {-6.99807816089155e+07, 1.23439533336424e+04, 7.41947367149584e+06, 5.14036819939642e+07,
                           -2.54666356840017e+13, 4.53268712277415e+09, 4.66829432137196e+12, 1.23144121825973e+14,
                           3.89994535976293e+12,  2.25670962111522e+16, -2.36027013243917e-04, -5.02039698946873e-04};
    const double ps[12] = {1.29148499223708e+00, 7.42273770077802e-01, 1.83956945157573e+00, 3.01474078952469e+00,
                           6.17493897637321e+00, 4.97136805912161e+00, 6.59152034486467e+00, 8.41614866661145e+00,
                           7.36125612301946e+00, 1.34956539961155e+01, 1.18675333603722e+00, 1.59183920812119e+00};
    const double r0 = 2.1, aa = 2.420229505061327e+07, bb = -1.036963443170083e+08, cc = 1.110675394653385e+08;
    const double x = xx < 2.5 ? 2.5 : xx;
    if (x <= r0) return aa * sq(x) + bb * x + cc;
    double s = pg[0] * std::exp(-ps[0] * x) / x;
    for (int k = 2; k <= 12; ++k) s = s + pg[k - 1] * ipow(x, k - 1) * std::exp(-ps[k - 1] * x);
    return s;
}

// alkali ion-He of Koutselos et al. (V_Koutselos_*_plus_gs :1701-1897)
struct IonFit { double v0, rho, Rm, C4, C6, C8; };
double koutselos(const IonFit& p, double xx) {
    const double a_bohr = 0.529177, hartree = 27.2114, ev_k = 11604.448;
    const double A = 146.98, a_a = 1.5024, B = 70.198, b_b = 1.4041;
    const double a_d = 1.3831, a_q = 2.4434, a_o = 10.614;
    const double CC4 = p.C4 + a_d / 2.0, CC6 = p.C6 + a_q / 2.0, CC8 = p.C8 + a_o / 2.0;
    const double x = xx / a_bohr, Rs = x / p.rho;
    const double vex = p.v0 * (A * std::exp(-a_a * Rs) - B * std::exp(-b_b * Rs));
    const double aux = p.Rm * 1.28;
    double h = 1.0;
    if (!(x >= aux)) h = x != 0.0 ? std::exp(-sq(aux / x - 1.0)) : 0.0;
    double s = vex;
    if (x != 0.0) {
        const double x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x6 * x2;
        s = s + (-(CC4 / x4) - CC6 / x6 - CC8 / x8) * h;
    }
    return s * hartree * ev_k;
}
const IonFit kCsPlus = {0.5092, 0.9241, 6.38, 0.0, 13.69, 236.4};
const IonFit kNaPlus = {0.2063, 0.7640, 4.55, 0.0, 1.424, 13.17};
const IonFit kRbPlus = {0.3474, 0.8868, 5.85, 0.0, 8.8, 126.2};
const IonFit kKPlus = {0.2690, 0.8589, 5.49, 0.0, 5.83, 72.05};

// surfaces (z in Angstrom): graphene (V_grafeno :2414, :2438), TiO2
// (V_TiO2 :2460), Au on TiO2 (:2498, :2513), Ag and Au on graphene
// (Ag_Graphene :936, dz_Ag_Graphene :953, Au_Graphene :3010, dz_Au_Graphene :3019)
double grafeno(double z) {
    const double mev_k = 11.604448, A = 554811.0, alpha = 3.21193, C4 = 1531.83, C6 = 22259.0;
    return (A * std::exp(-alpha * z) - C4 / cpow<4>(z) - C6 / cpow<6>(z)) * mev_k;
}
double grafeno_morse(double z) {
    const double mev_k = 11.604448, D = 0.493191, alpha = 1.54605, Ze = 4.34871;
    return (D * sq(1.0 - std::exp(-alpha * (z - Ze))) - D) * mev_k;
}
double tio2(double z) {
    const double cm_k = 1.4387770, de = 95.9583, alpha = 1.47089, ze = 4.22534, z0 = 4.77765, c3 = 8.85163,
                 aa = 1.06161;
    const double vz = de * sq(1.0 - std::exp(-alpha * (z - ze))) - de;
    const double vzd = -0.5 * (1.0 + std::tanh((z - z0) / aa) * c3 * cpow<3>(c3 / z));
    return (vz + vzd) * cm_k;
}
double au(double rr) {
    const double a1 = 1.6417, a2 = -0.6164, a3 = 0.3770, a4 = -0.0549, a5 = 0.0018, a6 = 0.0046, a7 = -0.0010,
                 a8 = 0.0001, a9 = 2.7e-6, Re = 4.124, De = 20.301143;
    const double r = rr - Re;
    return -De * (1.0 + r * (a1 + r * (a2 + r * (a3 + r * (a4 + r * (a5 + r * (a6 + r * (a7 + r * (a8 + r * a9))))))))) *
           std::exp(-a1 * r);
}
double au_tio2(double z, bool deriv) {
    const double C6 = -4.18518e6, C8 = 6.54459e7, C10 = -3.73112e8, C12 = 9.44942e8, C14 = -8.97265e8;
    const double mev_k = 11.604448;
    if (!(z > 2.0)) return (deriv ? -5415.35400390625 : 1821.44287109375) * mev_k;
    const double a = 1.0 / z, a2 = sq(a);
    if (!deriv) return (((((C14 * a2 + C12) * a2 + C10) * a2 + C8) * a2 + C6) * cpow<3>(a2)) * mev_k;
    return (((((-(14.0 * C14 * a2) - 12.0 * C12) * a2 - 10.0 * C10) * a2 - 8.0 * C8) * a2 - 6.0 * C6) *
            (cpow<3>(a2) * a)) * mev_k;
}
double on_graphene(double A, double beta, double C4, double unit, bool deriv, double z) {
    if (!deriv) return (A * std::exp(-beta * z) - C4 / cpow<4>(z)) * unit;
    return (-(A * beta * std::exp(-beta * z)) + 4.0 * C4 / cpow<5>(z)) * unit;
}

// CH3I-He Legendre components l = 0..4 (V_CH3I_l :3076-3162, radpot :3163):
// Morse d (1 - exp(-b (r - c)))^2 - d in atomic units, with each component's
// own inner cut-off and plateau
double ch3i(int l, double r) {
    const double dd[5] = {(double)0.000185445f, (double)-4.51657e-05f, (double)2.53688e-05f, (double)-6.51825e-06f,
                          (double)1.5032e-06f};
    const double bb[5] = {(double)0.782913f, (double)0.837569f, (double)0.849398f, (double)0.779254f,
                          (double)0.787553f};
    const double cc[5] = {(double)9.5243f, (double)10.3772f, (double)10.6988f, (double)11.454f, (double)12.236f};
    const double rcut[5] = {3.6, (double)3.6f, 3.6, 3.6, (double)3.6f};
    const double umax[5] = {3.1648655845e3, -5.1115736616e3, 5.5581502684e3, -2.7398540818e3, 2.4025334047e3};
    const double a_bohr = 0.529177, hartree = 27.2114, ev_k = 11604.448;
    if (r <= rcut[l]) return umax[l];
    const double ra = r / a_bohr;
    const double rp = dd[l] * sq(1.0 - std::exp(-bb[l] * (ra - cc[l]))) - dd[l];
    return rp * hartree * ev_k;
}

std::string trimmed(const char* s) {
    std::string t(s);
    while (!t.empty() && t.back() == ' ') t.pop_back();
    return t;
}

// V(r) of a closed-form potential; false when the name is not one of them
bool closed_form(const std::string& s, double r, double* v) {
    static const char* kAlk[5] = {"LI", "NA", "K", "RB", "CS"};
    for (int e = 0; e < 5; ++e)
        if (s == kAlk[e]) { *v = alkali(e, r); return true; }
    if (s == "Xe_He") *v = tang_toennies(74.9002, 1.8175, 21.5674, 430.148, 10610.21, r);
    else if (s == "Ar_He") *v = tang_toennies(75.8953, 2.0589, 9.7697, 152.8468, 3249.728, r);
    else if (s == "Aziz_He") *v = aziz_he(r);
    else if (s == "Ag_gs") *v = ag_fit(kAgGs, r);
    else if (s == "Ag_Pi") *v = ag_fit(kAgPi, r);
    else if (s == "Ag_Sig") *v = ag_fit(kAgSig, r);
    else if (s == "Ba_gs") *v = ba_gs(r);
    else if (s == "Ba_plus_old_gs") *v = ba_plus_old_gs(r);
    else if (s == "Cs_7s") *v = cs_7s(r);
    else if (s == "Cs_plus_Koutselos") *v = koutselos(kCsPlus, r);
    else if (s == "Na_plus_Koutselos") *v = koutselos(kNaPlus, r);
    else if (s == "Rb_plus_Koutselos") *v = koutselos(kRbPlus, r);
    else if (s == "K_plus_Koutselos") *v = koutselos(kKPlus, r);
    else if (s == "Grafeno") *v = grafeno(r);
    else if (s == "Grafeno_sin_dispersion") *v = grafeno_morse(r);
    else if (s == "TiO2") *v = tio2(r);
    else if (s == "Au") *v = au(r);
    else if (s == "Au_TiO2") *v = au_tio2(r, false);
    else if (s == "dAu_TiO2") *v = au_tio2(r, true);
    else if (s == "Ag_Graphene") *v = on_graphene(11557.3, 2.33671, 1143.0, 5.03216592455e2, false, r);
    else if (s == "dz_Ag_Graphene") *v = on_graphene(11557.3, 2.33671, 1143.0, 5.03216592455e2, true, r);
    else if (s == "Au_Graphene") *v = on_graphene(448058.0, 2.12336, 83077.3, 11.604448, false, r);
    else if (s == "dz_Au_Graphene") *v = on_graphene(448058.0, 2.12336, 83077.3, 11.604448, true, r);
    else if (s.size() == 6 && s.compare(0, 5, "CH3I_") == 0 && s[5] >= '0' && s[5] <= '4') *v = ch3i(s[5] - '0', r);
    else return false;
    return true;
}

}  // namespace

bool select_pot(const char* selec, double r, double r_cutoff, double umax, double* v) {
    const std::string s = trimmed(selec);
    if (s == "null") { *v = 0.0; return true; }
    double val = 0.0;
    bool known = false;
    const bool cut = r < r_cutoff;
    for (const InvPowFit& f : kFits)
        if (s == f.selec) { known = true; if (!cut) val = inv_pow_fit(f, r); break; }
    if (!known) {
        double tmp;
        known = closed_form(s, 1.0, &tmp);
        if (known && !cut) closed_form(s, r, &val);
    }
    if (!known) return false;
    *v = cut ? umax : val;
    return true;
}

const char* known_potentials() {
    static std::string names;
    if (names.empty()) {
        names = "null LI NA K RB CS Xe_He Ar_He Aziz_He Ag_gs Ag_Pi Ag_Sig Ba_gs Ba_plus_old_gs Cs_7s "
                "Cs_plus_Koutselos Na_plus_Koutselos Rb_plus_Koutselos K_plus_Koutselos Grafeno "
                "Grafeno_sin_dispersion TiO2 Au Au_TiO2 dAu_TiO2 Ag_Graphene dz_Ag_Graphene Au_Graphene "
                "dz_Au_Graphene CH3I_0 CH3I_1 CH3I_2 CH3I_3 CH3I_4";
        for (const InvPowFit& f : kFits) {
            names += ' ';
            names += f.selec;
        }
    }
    return names.c_str();
}

// He on TiO2, laterally averaged: De (e^{2 alpha (Re - z)} - 2 e^{alpha (Re - z)})
// in cm^-1, times 1.4387752 K/cm^-1 written as a default-real literal
double morse_tio2_1d(double z) {
    const double De = 39.95, alpha = 1.721, Re = 4.3619;
    return De * (std::exp(2.0 * alpha * (Re - z)) - 2.0 * std::exp(alpha * (Re - z))) * (double)1.4387752f;
}

// He on TiO2 with the surface corrugation: De, alpha and Ze are each
// Lx^T M Ly with Lx = cos(2 pi [0,1,2] x/a), Ly = cos(2 pi [0,1,2,3] y/b)
// (products summed in the order of the matmul calls); K
double morse_tio2_3d(double x, double y, double z) {
    static const double kDe[3][4] = {{54.671500, -2.228480, 0.585194, 13.211000},
                                     {4.842040, 0.363525, 4.614530, -1.182030},
                                     {3.747220, 3.222690, 0.333350, 0.587320}};
    static const double kAlpha[3][4] = {{1.492420, -0.010921, 0.003189, -0.168927},
                                        {-0.097618, -0.024780, 0.062441, 0.016801},
                                        {-0.034535, 0.030015, 0.012846, -0.011699}};
    static const double kZe[3][4] = {{4.073990, 0.057537, -0.001500, -0.568990},
                                     {-0.115578, -0.018127, -0.090046, 0.008590},
                                     {-0.075153, -0.051210, -0.023349, -0.017931}};
    const double pi = std::acos(-1.0), a = 2.9850, b = 6.49548;
    double lx[3], ly[4];
    for (int i = 0; i < 3; ++i) lx[i] = std::cos(2.0 * pi * (double)i * x / a);
    for (int j = 0; j < 4; ++j) ly[j] = std::cos(2.0 * pi * (double)j * y / b);
    auto form = [&](const double (*M)[4]) {
        double s = 0.0;
        for (int j = 0; j < 4; ++j) {
            double t = 0.0;
            for (int i = 0; i < 3; ++i) t = t + lx[i] * M[i][j];
            s = s + t * ly[j];
        }
        return s;
    };
    const double De = form(kDe), al = form(kAlpha), Ze = form(kZe);
    return De * (std::exp(-(2.0 * al * (z - Ze))) - 2.0 * std::exp(-(al * (z - Ze))));
}

}  // namespace imp
}  // namespace hetd
