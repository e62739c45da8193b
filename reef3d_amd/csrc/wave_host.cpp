// Host-side set-up of relaxation-zone wave generation (iowave, B98 = 2) and the numerical
// beach (B99 = 1, 2) of one tank: the linear-wave parameters, the zone polygons, and per
// column the zone membership, the relaxation factors and the distances along the wave.
// Everything here is static for a run, so it is computed once in FP64 on the host with the
// C library (the same functions the reference calls) and uploaded; the kernels only
// evaluate the time-dependent wave at the cells of the zones.
#include "wave.h"

#include <cmath>

namespace reef {

namespace {

// iowave::intriangle (iowave_dist_ini.cpp:141-190): is (x0, y0) in the triangle A B C,
// by the vertical ray through the point; the test's || / && precedence is the reference's
int intriangle(double Ax, double Ay, double Bx, double By, double Cx, double Cy, double x0, double y0)
{
    const double Px = x0, Py = y0, Pz = -10.0;
    const double Qx = x0, Qy = y0, Qz = +10.0;
    const double PQx = Qx - Px, PQy = Qy - Py, PQz = Qz - Pz;
    const double Mx = PQy * Pz - PQz * Py;
    const double My = PQz * Px - PQx * Pz;
    const double u = PQz * (Cx * By - Cy * Bx) + Mx * (Cx - Bx) + My * (Cy - By);
    const double v = PQz * (Ax * Cy - Ay * Cx) + Mx * (Ax - Cx) + My * (Ay - Cy);
    const double w = PQz * (Bx * Ay - By * Ax) + Mx * (Bx - Ax) + My * (By - Ay);
    int check = 1;
    if(u == 0.0 && v == 0.0 && w == 0.0) check = 0;
    if((u >= 0.0 && v >= 0.0 && w >= 0.0) || ((u < 0.0 && v < 0.0 && w < 0.0) && check == 1)) return 1;
    return 0;
}

// one relaxation zone: the line from S to E and the four corners of the band of half
// width d around it (iowave::distgen_ini / distbeach_ini, iowave_dist_ini.cpp:27-139)
struct Zone
{
    double S[2], E[2], C1[2], C2[2], C3[2], C4[2], d;

    Zone(double xs, double ys, double xe, double ye, double dist)
    {
        S[0] = xs;
        S[1] = ys;
        E[0] = xe;
        E[1] = ye;
        d = dist;
        const double dx = E[0] - S[0], dy = E[1] - S[1];
        const double n1x = -dy, n1y = dx, n2x = dy, n2y = -dx;
        double l1 = std::sqrt(std::pow(n1x, 2.0) + std::pow(n1y, 2.0));
        double l2 = std::sqrt(std::pow(n2x, 2.0) + std::pow(n2y, 2.0));
        l1 = std::fabs(l1) > 1.0e-20 ? l1 : 1.0e20;
        l2 = std::fabs(l2) > 1.0e-20 ? l2 : 1.0e20;
        C1[0] = S[0] + (n1x / l1) * dist;
        C1[1] = S[1] + (n1y / l1) * dist;
        C2[0] = S[0] + (n2x / l2) * dist;
        C2[1] = S[1] + (n2y / l2) * dist;
        C3[0] = E[0] + (n1x / l1) * dist;
        C3[1] = E[1] + (n1y / l1) * dist;
        C4[0] = E[0] + (n2x / l2) * dist;
        C4[1] = E[1] + (n2y / l2) * dist;
    }
    // the two triangles C1 C3 C2 and C3 C4 C2 of the band
    bool inside(double x0, double y0) const
    {
        return intriangle(C1[0], C1[1], C3[0], C3[1], C2[0], C2[1], x0, y0) == 1
            || intriangle(C3[0], C3[1], C4[0], C4[1], C2[0], C2[1], x0, y0) == 1;
    }
    // distance of (x0, y0) to the line (the distgen / distbeach formula)
    double dist(double x0, double y0) const
    {
        double denom = std::sqrt(std::pow(E[1] - S[1], 2.0) + std::pow(E[0] - S[0], 2.0));
        denom = denom > 1.0e-20 ? denom : 1.0e20;
        return std::fabs((E[1] - S[1]) * x0 - (E[0] - S[0]) * y0 + E[0] * S[1] - E[1] * S[0]) / denom;
    }
};

const double EE_REF = 2.71828182846;   // definitions.h:38

// rb1_ext (iowave_relax_functions.cpp:28-95) for one generation zone: 1 outside it, inside
// 1 - (exp(x^3.5) - 1) / (EE - 1) with x = max(1 - dist / d, 0)
double relax_gen(const Zone& z, double x0, double y0)
{
    if(!z.inside(x0, y0)) return 1.0;
    double x = -1.0e20;
    const double xdist = std::fmin(z.dist(x0, y0), 1.0e20);
    x = std::fmax(1.0 - xdist / z.d, x);
    x = std::fmax(x, 0.0);
    return 1.0 - (std::exp(std::pow(x, 3.5)) - 1.0) / (EE_REF - 1.0);
}

// rb3_ext (iowave_relax_functions.cpp:143-200) for one beach zone: x = max((d - dist) /
// (d * dist2_fac), 0), the same function of x, averaged over the zones that contain the
// point (one here)
double relax_beach(const Zone& z, double x0, double y0, double fac)
{
    if(!z.inside(x0, y0)) return 1.0;
    double x = std::fmin(z.dist(x0, y0), 1.0e20);
    x = (z.d - std::fabs(x)) / (z.d * fac);
    x = std::fmax(x, 0.0);
    double r = 0.0;
    r += 1.0 - (std::exp(std::pow(x, 3.5)) - 1.0) / (EE_REF - 1.0);
    return r / double(1);
}

}  // namespace

std::string build_wave(const HostGrid& h, const reef_wave_desc& w, WaveSetup& out)
{
    if(w.B98 != 2 && w.B98 != 0) return "wave generation B 98: 2 (relaxation) or 0 (none) only";
    if(w.B99 != 0 && w.B99 != 1 && w.B99 != 2) return "numerical beach B 99: 0, 1 or 2 only";
    if(w.B92 != 2) return "wave theory B 92: 2 (linear waves, wave_lib_linear) only";
    if((w.B91 == 1) == (w.B93 == 1)) return "wave parameters: give either B 91 (height, length) or B 93 (height, period)";
    if(w.B101 < 0 || w.B101 > 2) return "ramp B 101: 0, 1 or 2";
    if(w.B98 == 2 && !(w.B96_1 > 0.0)) return "generation zone length B 96 must be > 0";
    if(w.B99 != 0 && !(w.B96_2 > 0.0)) return "beach zone length B 96 must be > 0";
    const double PI = PI_REF;
    WaveDev& d = out.dev;
    d = WaveDev{};
    // wave_lib_parameters (wave_lib_parameters.cpp:28-160) for B92 = 2, water depth phimean
    // (B94 = 0)
    const double phimean = w.F60;
    const double wdt = phimean;
    double wa = 0.0, wH = 0.0, wL = 0.0, wk = 0.0, ww = 0.0, wf = 0.0, wT = 0.0;
    if(w.B91 == 1)
    {
        wa = 0.5 * w.B91_1;
        wH = w.B91_1;
        wL = w.B91_2;
        wk = (2.0 * PI) / (wL > 1.0e-20 ? wL : 1.0e20);
        ww = std::sqrt(std::fabs(9.81 * wk * std::tanh(wk * (wdt))));
        wf = ww / (2.0 * PI);
        wT = 1.0 / wf;
    }
    else
    {
        wa = 0.5 * w.B93_1;
        wH = w.B93_1;
        wT = w.B93_2;
        const double wL0 = (9.81 / (2.0 * PI)) * wT * wT;
        const double k0 = (2.0 * PI) / wL0;
        const double S0 = std::sqrt(k0 * wdt) * (1.0 + (k0 * wdt) / 6.0 + (k0 * k0 * wdt * wdt) / 30.0);
        wL = wL0 * std::tanh(S0);
        for(int qn = 0; qn < 500; ++qn) wL = wL0 * std::tanh(2.0 * PI * wdt / wL);
        wf = 1.0 / wT;
        ww = wf * 2.0 * PI;
        wk = (2.0 * PI) / (wL > 1.0e-20 ? wL : 1.0e20);
    }
    d.wa = wa;
    d.wk = wk;
    d.ww = ww;
    d.wdt = wdt;
    d.wD = phimean;                                      // wave_interface.cpp:55-59
    d.pshift = w.B120 * (PI / 180.0);
    d.cosg = std::cos((w.B105_1) * (PI / 180.0));        // wave_lib_linear.cpp:37-38
    d.sing = std::sin((w.B105_1) * (PI / 180.0));
    d.phimean = phimean;
    d.psi = 0.6 * h.DXM;                                 // iowave.cpp:32
    d.epsi = 3.0 * h.DXM;
    d.W50 = w.W50;
    d.Ui = 0.0;                                          // W10 = 0 and no inflow cells (Qin)
    d.gen = w.B98 == 2;
    d.beach = w.B99 == 1 || w.B99 == 2;
    d.beach_phi = w.B99 == 2;
    d.ramp_mode = w.B101;
    d.ramp_T = w.B101 == 1 ? w.B102 * wT : w.B102;
    out.wT = wT;
    out.wL = wL;
    out.wH = wH;

    // zones (iowave.cpp:96-150): the generation line along x = xmin, the beach line along
    // x = xmax, each reaching 10 DXM beyond the domain in y, half width B96_1 / B96_2
    const int nx = h.nx, ny = h.ny;
    const double* XN = h.X[0].data();
    const double* YN = h.X[1].data();
    const double* XP = h.P[0].data();
    const double* YP = h.P[1].data();
    const double xmin = XN[MARGE], xmax = XN[nx + MARGE], ymin = YN[MARGE], ymax = YN[ny + MARGE];
    const Zone gen(xmin, ymin - 10.0 * h.DXM, xmin, ymax + 10.0 * h.DXM, w.B96_1);
    const Zone beach(xmax, ymin - 10.0 * h.DXM, xmax, ymax + 10.0 * h.DXM, w.B96_2);
    const double dist2_fac = w.B99 == 1 ? 2.0 : 1.0;     // iowave.cpp:38-41
    // the wave's origin (B105 = 0: the domain's lower corner) and direction (iowave.cpp:151-158)
    const double x0 = xmin, y0 = ymin;
    const double tan_alpha = std::tan((w.B105_1 + 90.0) * (PI / 180.0));
    auto xgen = [&](double x1, double y1) {
        return std::fabs(y1 - tan_alpha * x1 + tan_alpha * x0 - y0) / std::sqrt(std::pow(tan_alpha, 2.0) + 1.0);
    };
    const size_t ncol = (size_t)nx * ny;
    out.col.assign(WC_N * ncol, 0.0);
    auto C = [&](int a, int i, int j) -> double& { return out.col[a * ncol + (size_t)i + (size_t)nx * j]; };
    for(int j = 0; j < ny; ++j)
        for(int i = 0; i < nx; ++i)
        {
            const int ip = i + MARGE, jp = j + MARGE;
            // distgen / distbeach at the cell centre (iowave_dist.cpp:130-190): < 1e20 inside
            C(WC_GEN, i, j) = (w.B98 == 2 && gen.inside(XP[ip], YP[jp])) ? 1.0 : 0.0;
            C(WC_BEACH, i, j) = (w.B99 != 0 && beach.inside(XP[ip], YP[jp])) ? 1.0 : 0.0;
            // wavegen_precalc_relax_func (iowave_precalc_relax_func.cpp:27-101): the u, v and
            // cell points (pos1, pos2, pos)
            const double x1 = XN[ip + 1], y1 = YP[jp];
            const double x2 = XP[ip], y2 = YN[jp + 1];
            const double x4 = XP[ip], y4 = YP[jp];
            C(WC_R1WG, i, j) = w.B98 == 2 ? relax_gen(gen, x1, y1) : 1.0;
            C(WC_R2WG, i, j) = w.B98 == 2 ? relax_gen(gen, x2, y2) : 1.0;
            C(WC_R4WG, i, j) = w.B98 == 2 ? relax_gen(gen, x4, y4) : 1.0;
            C(WC_R1NB, i, j) = w.B99 != 0 ? relax_beach(beach, x1, y1, dist2_fac) : 1.0;
            C(WC_R2NB, i, j) = w.B99 != 0 ? relax_beach(beach, x2, y2, dist2_fac) : 1.0;
            C(WC_R4NB, i, j) = w.B99 != 0 ? relax_beach(beach, x4, y4, dist2_fac) : 1.0;
            // xgen, xgen1, xgen2 (iowave_dist.cpp:27-75)
            C(WC_XG4, i, j) = xgen(x4, y4);
            C(WC_XG1, i, j) = xgen(x1, y1);
            C(WC_XG2, i, j) = xgen(x2, y2);
        }
    return std::string();
}

}  // namespace reef
