// CPU oracle for the REEF3D free-surface hot path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker of tests/ and the CPU baseline timed by
// bench.py.  The product (reef3d_amd/) never includes, links or calls this file.
//
// A plain, serial restatement of the reference's algorithm, in the reference's loop order
// and operation order (built with -ffp-contract=off) so that it reproduces the reference
// bit for bit on the same inputs.  Every function cites the reference file:line it follows
// (paths relative to REEF3D src/).  Box domains only: every side is a wall (group 21) or
// the top is a free-surface boundary (group 3); one rank; no solids, porosity 1.
#include "reef_oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

const double PI_REF = 3.14159265359;   // definitions.h:37
const int MARGE = 5;                    // 1-D array offset (iterators1D.h)
const int OBJ = -20, WATER = 10;        // definitions.h:27-34

struct Entry
{
    int i, j, k, side, group;
    double dist;
};

struct Dgc
{
    int i, j, k, di, dj, dk;
};

}  // namespace

struct orc_domain
{
    int nx, ny, nz, jdir;
    int F49 = 1;                // reinit discretisation: 1 reinidisc_f, 0 reinidisc_fsf
    int D20 = 0;                // momentum diffusion: 0 diff_void, 1 ediff2, 2 idiff2_FS (implicit)
    double N43 = 1.0e-5;        // stop criterion of the diffusion solves (ini.cpp:589)
    int df_set = 0;             // direct-forcing flags DF1..3 set (after the first forcing call)
    // private x of the solver of the diffusion solves (psolv, one bicgstab_ijk object for
    // var 1..3, driver_logic_cfd.cpp:464-468): fillxvec writes only the FLEXLOOP cells of
    // each call, so the other cells keep what earlier calls of other vars left there and
    // the matvec of a restart reads them (bicgstab_ijk.cpp:280-294, 366-378)
    std::vector<double> psolv_x;
    int dot_ext = 0;            // 1: BiCGStab dot products accumulated in long double
    // threads of the per-cell loops (stencils, stage combinations, BiCGStab sweeps): 1 keeps
    // every result bit-exact with the reference; more keep the per-cell results and sum the
    // BiCGStab dot products per i slab (a different rounding); volume sums stay serial
    int threads = 1;
    // iowave relaxation-zone wave generation (B98 = 2) and numerical beach (B99), linear
    // waves; on = 0: the void ioflow of the other cases
    struct Wave
    {
        int on = 0;
        orc_wave_desc w;
        double wa, wk, ww, wdt, wT, wL, pshift, singamma, cosgamma, phimean, psi, epsi, tan_alpha, x0, y0;
        double G[4][2], Gs[2], Ge[2], B[4][2], Bs[2], Be[2];   // zone corners and lines
        double dist2_fac;
        std::vector<double> eta;                // slice4 eta (i, j), zero until written
        std::vector<double> uval, vval, wval, lsval;   // per cell of the precalc
    } wave;
    double simtime = 0.0;       // p->simtime of the wave runs
    long sx, sy, sz;
    int group[7];
    orc_params prm;
    std::vector<double> X[3];     // node coordinates XN/YN/ZN, offset MARGE, zero padded
    std::vector<double> P[3];     // XP/YP/ZP
    std::vector<double> DN[3];    // DXN/DYN/DZN
    std::vector<double> DP[3];    // DXP/DYP/DZP
    double DXM, DYM, DZM, psi, ydir;
    long cellnum;
    // weno tables [IP][uf][row][coef] (weno_nug_func.h:198-200)
    std::vector<double> qf[3], cf[3], isf[3];
    // boundary lists
    std::vector<Entry> gcb[5];
    std::vector<Dgc> dgc[5];
    std::vector<int> flag[5];     // flag1..flag4 over the whole field (ghosts included)
    int last[4];                  // ulast, vlast, wlast

    long id(int i, int j, int k) const { return (long)(i + 3) + sx * ((long)(j + 3) + sy * (long)(k + 3)); }
    int nn(int a) const { return a == 0 ? nx : a == 1 ? ny : nz; }
};

namespace {

// ------------------------------------------------------------------ grid
// grid::gridspacing (grid.cpp:54-150); DXM averages DXP over x, DYP over y (3D), DZP over z.
void gridspacing(orc_domain* d)
{
    for(int a = 0; a < 3; ++a)
    {
        int n = d->nn(a);
        int len = n + 1 + 4 * MARGE;
        std::vector<double>& X = d->X[a];
        d->P[a].assign(len, 0.0);
        d->DN[a].assign(len, 0.0);
        d->DP[a].assign(len, 0.0);
        for(int i = -MARGE; i < n + MARGE; ++i)
        {
            int ip = i + MARGE;
            d->P[a][ip] = 0.5 * (X[ip] + X[ip + 1]);
            d->DN[a][ip] = X[ip + 1] - X[ip];
            d->DP[a][ip] = 0.5 * (X[ip + 2] + X[ip + 1]) - 0.5 * (X[ip + 1] + X[ip]);
        }
    }
    double DXM = 0.0;
    int count = 0;
    for(int i = 0; i < d->nx; ++i) { DXM += d->DP[0][i + MARGE]; ++count; }
    if(d->jdir == 1)
        for(int j = 0; j < d->ny; ++j) { DXM += d->DP[1][j + MARGE]; ++count; }
    for(int k = 0; k < d->nz; ++k) { DXM += d->DP[2][k + MARGE]; ++count; }
    d->DXM = DXM / double(count);
}

// weno_nug_func_precalc_qf.cpp:27-322, _cf.cpp:27-188, _isf.cpp:27-890 restated for one
// coordinate array X (XN for uf=0, XP for uf=1) and one index IP.
void tables_at(const double* X, int ip, double* q /*6x2*/, double* c /*6*/, double* s /*6x3*/)
{
    const double m2 = X[ip - 2], m1 = X[ip - 1], x0 = X[ip], p1 = X[ip + 1], p2 = X[ip + 2],
                 p3 = X[ip + 3], p4 = X[ip + 4];
    // qf: reconstruction coefficients
    q[0] = ((p2 - p1) / (p3 - x0)) * ((p3 - p1) / (p2 - x0));
    q[1] = ((p2 - p1) / (p3 - x0)) * ((p1 - x0) / (p3 - p1));
    q[2] = ((p1 - x0) / (p2 - m1)) * ((p1 - m1) / (p2 - x0));
    q[3] = ((p1 - x0) / (p2 - m1)) * ((p2 - p1) / (p1 - m1));
    q[4] = ((p1 - x0) / (x0 - m2)) * ((p1 - m1) / (p1 - m2));
    q[5] = (1.0 + (p1 - x0) / (p1 - m1) + (p1 - x0) / (p1 - m2));
    q[6] = (1.0 + (p2 - p1) / (p3 - p1) + (p2 - p1) / (p4 - p1));
    q[7] = ((p2 - p1) / (p4 - p1)) * ((p3 - p1) / (p4 - p2));
    q[8] = ((p2 - p1) / (p3 - x0)) * ((p3 - p1) / (p2 - x0));
    q[9] = ((p2 - p1) / (p3 - x0)) * ((p1 - x0) / (p3 - p1));
    q[10] = ((p1 - x0) / (p2 - m1)) * ((p1 - m1) / (p2 - x0));
    q[11] = ((p1 - x0) / (p2 - m1)) * ((p2 - p1) / (p1 - m1));
    // cf: linear weights
    c[0] = ((p1 - m2) / (p3 - m2)) * ((p1 - m1) / (p3 - m1));
    c[1] = ((p1 - m2) / (p3 - m2)) * ((p3 - p1) / (p3 - m1)) * ((p3 - m1) / (p2 - m2) + 1.0);
    c[2] = ((p2 - p1) / (p3 - m2)) * ((p3 - p1) / (p2 - m2));
    c[3] = ((p1 - x0) / (p4 - m1)) * ((p1 - m1) / (p4 - x0));
    c[4] = ((p1 - m1) / (p4 - m1)) * ((p4 - p1) / (p4 - x0)) * ((p4 - x0) / (p3 - m1) + 1.0);
    c[5] = ((p3 - p1) / (p4 - m1)) * ((p4 - p1) / (p3 - m1));
    // isf: smoothness indicator coefficients
    double fac;
    fac = 4.0 * pow((p1 - x0) / (p3 - x0), 2.0);
    s[0] = fac * ((10.0 * pow(p1 - x0, 2.0) + (p2 - x0) * (p2 - p1)) / pow(p3 - p1, 2.0));
    s[1] = fac * ((20.0 * pow(p1 - x0, 2.0) + 2.0 * (p2 - x0) * (p2 - p1) + (p3 - x0) * (2.0 * p2 - p1 - x0))
                  / ((p3 - p1) * (p2 - x0)));
    s[2] = fac * ((10.0 * pow(p1 - x0, 2.0) + (p3 + p2 - 2.0 * x0) * (p3 + p2 - p1 - x0)) / pow(p2 - x0, 2.0));
    fac = 4.0 * pow((p1 - x0) / (p2 - m1), 2.0);
    s[3] = fac * ((10.0 * pow(p1 - x0, 2.0) + (p2 - x0) * (p2 - p1)) / pow(p1 - m1, 2.0));
    s[4] = fac * ((20.0 * pow(p1 - x0, 2.0) - (p2 - p1) * (x0 - m1) - (p2 - x0) * (p1 - m1))
                  / ((p2 - x0) * (p1 - m1)));
    s[5] = fac * ((10.0 * pow(p1 - x0, 2.0) + (x0 - m1) * (p1 - m1)) / pow(p2 - x0, 2.0));
    fac = 4.0 * pow((p1 - x0) / (p1 - m2), 2.0);
    s[6] = fac * ((10.0 * pow(p1 - x0, 2.0) + (p1 - m1) * (x0 - m1)) / pow(x0 - m2, 2.0));
    s[7] = fac * ((20.0 * pow(p1 - x0, 2.0) + 2.0 * (p1 - m1) * (x0 - m1) + (p1 - m2) * (p1 + x0 - 2.0 * m1))
                  / ((p1 - m1) * (x0 - m2)));
    s[8] = fac * ((10.0 * pow(p1 - x0, 2.0) + (2.0 * p1 - m2 - m1) * (p1 + x0 - m1 - m2)) / pow(p1 - m1, 2.0));
    fac = 4.0 * pow((p2 - p1) / (p4 - p1), 2.0);
    s[9] = fac * ((10.0 * pow(p2 - p1, 2.0) + (p3 - p1) * (p3 - p2)) / pow(p4 - p2, 2.0));
    s[10] = fac * ((20.0 * pow(p2 - p1, 2.0) + 2.0 * (p3 - p1) * (p3 - p2) + (p4 - p1) * (2.0 * p3 - p2 - p1))
                   / ((p4 - p2) * (p3 - p1)));
    s[11] = fac * ((10.0 * pow(p2 - p1, 2.0) + (p4 + p3 - 2.0 * p1) * (p4 + p3 - p2 - p1)) / pow(p3 - p1, 2.0));
    fac = 4.0 * pow((p2 - p1) / (p3 - x0), 2.0);
    s[12] = fac * ((10.0 * pow(p2 - p1, 2.0) + (p3 - p1) * (p3 - p2)) / pow(p2 - x0, 2.0));
    s[13] = fac * ((20.0 * pow(p2 - p1, 2.0) - (p3 - p2) * (p1 - x0) - (p3 - p1) * (p2 - x0))
                   / ((p3 - p1) * (p2 - x0)));
    s[14] = fac * ((10.0 * pow(p2 - p1, 2.0) + (p1 - x0) * (p2 - x0)) / pow(p3 - p1, 2.0));
    fac = 4.0 * pow((p2 - p1) / (p2 - m1), 2.0);
    s[15] = fac * ((10.0 * pow(p2 - p1, 2.0) + (p2 - x0) * (p1 - x0)) / pow(p1 - m1, 2.0));
    s[16] = fac * ((20.0 * pow(p2 - p1, 2.0) + 2.0 * (p2 - x0) * (p1 - x0) + (p2 - m1) * (p2 + p1 - 2.0 * x0))
                   / ((p2 - x0) * (p1 - m1)));
    s[17] = fac * ((10.0 * pow(p2 - p1, 2.0) + (2.0 * p2 - m1 - x0) * (p2 + p1 - x0 - m1)) / pow(p2 - x0, 2.0));
}

void build_tables(orc_domain* d)
{
    for(int a = 0; a < 3; ++a)
    {
        int n = d->nn(a);
        d->qf[a].assign((size_t)(n + 8) * 24, 0.0);
        d->cf[a].assign((size_t)(n + 8) * 12, 0.0);
        d->isf[a].assign((size_t)(n + 8) * 36, 0.0);
        for(int uf = 0; uf < 2; ++uf)
        {
            const double* X = uf == 0 ? d->X[a].data() : d->P[a].data();
            for(int i = -1; i < n + 1; ++i)      // IBLOOP
            {
                int ip = i + MARGE;
                tables_at(X, ip, &d->qf[a][(size_t)ip * 24 + uf * 12], &d->cf[a][(size_t)ip * 12 + uf * 6],
                          &d->isf[a][(size_t)ip * 36 + uf * 18]);
            }
        }
    }
}

// ------------------------------------------------------------------ boundary lists
// read_grid.cpp:458-488 (gcb4 as written for a box), lexer_gridini.cpp:146-166 (gcd4),
// gc_flagfield.cpp:26-98 (flags), grid_helper_gcb1/2/3.cpp (velocity lists),
// grid_helper.cpp:42-59 + grid_helper_dgc1..4.cpp (edge/corner lists; the hgc marks
// accumulate from dgc1 to dgc4 because make_dgc clears them once).
void build_lists(orc_domain* d)
{
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    long ncell = d->sx * d->sy * d->sz;
    // gcb4 in LOOP order, sides 1..6
    std::vector<Entry>& g4 = d->gcb[4];
    g4.clear();
    for(int i = 0; i < nx; ++i)
    for(int j = 0; j < ny; ++j)
    for(int k = 0; k < nz; ++k)
    {
        int at[7] = {0, i == 0, j == ny - 1, j == 0, i == nx - 1, k == 0, k == nz - 1};
        for(int s = 1; s <= 6; ++s)
            if(at[s])
            {
                Entry e{i, j, k, s, d->group[s], 0.0};
                if(s == 1 || s == 4) e.dist = 0.5 * d->DP[0][i + MARGE];
                else if(s == 2 || s == 3) e.dist = 0.5 * d->DP[1][j + MARGE];
                else e.dist = 0.5 * d->DP[2][k + MARGE];
                g4.push_back(e);
            }
    }
    // flags
    for(int f = 1; f <= 4; ++f)
        d->flag[f].assign(ncell, OBJ);
    for(int i = 0; i < nx; ++i)
    for(int j = 0; j < ny; ++j)
    for(int k = 0; k < nz; ++k)
        for(int f = 1; f <= 4; ++f)
            d->flag[f][d->id(i, j, k)] = WATER;
    for(const Entry& e : g4)
    {
        if(e.side == 4) d->flag[1][d->id(e.i, e.j, e.k)] = OBJ;
        if(e.side == 2) d->flag[2][d->id(e.i, e.j, e.k)] = OBJ;
        if(e.side == 6) d->flag[3][d->id(e.i, e.j, e.k)] = OBJ;
    }
    // velocity lists (fillgcb1/2/3)
    const int lastside[4] = {0, 4, 2, 6};
    for(int c = 1; c <= 3; ++c)
    {
        std::vector<Entry>& g = d->gcb[c];
        g = g4;
        int ax = c - 1;
        std::vector<char> fgc(ncell, 0);     // fgc[IJK][lastside-1]
        for(Entry& e : g)
        {
            int sa = (e.side == 1 || e.side == 4) ? 0 : (e.side == 2 || e.side == 3) ? 1 : 2;
            if(sa == ax)
            {
                int idx = ax == 0 ? e.i : ax == 1 ? e.j : e.k;
                e.dist += 0.5 * d->DP[ax][idx + MARGE];
            }
            if(e.side == lastside[c]) fgc[d->id(e.i, e.j, e.k)] = 1;
        }
        for(Entry& e : g)
            if(e.side == lastside[c])
            {
                if(c == 1) e.i -= 1;
                if(c == 2) e.j -= 1;
                if(c == 3) e.k -= 1;
            }
        for(Entry& e : g)
            if(e.side != lastside[c] && fgc[d->id(e.i, e.j, e.k)] == 1)
                e.side = -std::abs(e.side);
    }
    // edge / corner lists
    static const int DIRS[20][3] = {
        {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0},
        {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
        {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1},
        {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
        {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
    std::vector<int> hgc(ncell, 0);
    for(int c = 1; c <= 4; ++c)
    {
        for(const Entry& e : d->gcb[c])
        {
            if(e.side <= 0) continue;
            for(int n = 0; n < 3; ++n)
            {
                int i = e.i, j = e.j, k = e.k;
                if(e.side == 1) i -= n + 1;
                if(e.side == 4) i += n + 1;
                if(e.side == 3) j -= n + 1;
                if(e.side == 2) j += n + 1;
                if(e.side == 5) k -= n + 1;
                if(e.side == 6) k += n + 1;
                hgc[d->id(i, j, k)] += 1;
            }
        }
        const std::vector<int>& fl = d->flag[c];
        int il = c == 1 ? d->last[1] : 0, jl = c == 2 ? d->last[2] : 0, kl = c == 3 ? d->last[3] : 0;
        std::vector<Dgc>& out = d->dgc[c];
        out.clear();
        for(int i = 0; i < nx - il; ++i)
        for(int j = 0; j < ny - jl; ++j)
        for(int k = 0; k < nz - kl; ++k)
        {
            if(fl[d->id(i, j, k)] <= 0) continue;
            for(int q = 0; q < 20; ++q)
            {
                int di = DIRS[q][0], dj = DIRS[q][1], dk = DIRS[q][2];
                if(fl[d->id(i + di, j + dj, k + dk)] >= 0) continue;
                if(di != 0 && fl[d->id(i + di, j, k)] >= 0) continue;
                if(dj != 0 && fl[d->id(i, j + dj, k)] >= 0) continue;
                if(dk != 0 && fl[d->id(i, j, k + dk)] >= 0) continue;
                if(hgc[d->id(i + di, j + dj, k + dk)] != 0) continue;
                out.push_back(Dgc{i, j, k, di, dj, dk});
            }
        }
    }
}

// ------------------------------------------------------------------ ghost cells
// ghostcell labels (ghostcell.cpp:45-200 with the defaults B20=2, B23=1, B75..B77=1,
// B98=B99=0) and gceval1..4 (gc_epol1.cpp:26-121, gc_epol2.cpp, gc_epol3.cpp,
// gc_epol4.cpp:26-317) for the groups of a box (21 wall, 3 free surface) and the field
// tags of the hot path.  Returns 0 (untouched), 1 (dirichlet_ortho), 4 (neumann),
// 5 (noslip).
int label(int kind, int gcv, int bc, int cs)
{
    if(gcv == 50) return 4;
    if(kind == 4)
    {
        if((gcv == 51 || gcv == 52 || gcv == 53 || gcv == 54) && (bc == 21 || bc == 3)) return 4;
        if(gcv == 40 && (bc == 21 || bc == 3)) return 4;
        if(gcv == 1) return 4;
        return 0;
    }
    // velocities: tangential / normal sides of component `kind`
    bool normal = (kind == 1 && (cs == 1 || cs == 4)) || (kind == 2 && (cs == 2 || cs == 3))
               || (kind == 3 && (cs == 5 || cs == 6));
    int g_main = kind == 1 ? 10 : kind == 2 ? 11 : 12;   // gcval_u/v/w
    int g_alt = kind;                                      // gcv 1,2,3 (same rules)
    int g_corr = kind == 1 ? 7 : kind == 2 ? 8 : 9;      // pjm_corr vel_setup
    int g_neu = kind == 1 ? 14 : kind == 2 ? 15 : 16;
    if(bc == 21)
    {
        if(!normal && (gcv == g_main || gcv == g_alt)) return 5;   // gclabel_u/v/w
        if(gcv == g_neu) return 4;
        if(normal && (gcv == g_main || gcv == g_alt)) return 1;    // gclabel_*_orth
        if(normal && gcv == g_corr) return 5;                        // gclabel_vel
        return 0;
    }
    if(bc == 3)
    {
        int g_fs = kind == 1 ? 17 : kind == 2 ? 18 : 19;
        if(gcv == g_main || gcv == g_fs || gcv == g_alt)
        {
            if(!normal) return 4;
            return kind == 3 ? 5 : 1;   // gc_epol3.cpp: free-surface w normal -> noslip (A10=6)
        }
        return 0;
    }
    return 0;
}

inline void setg(orc_domain* d, double* f, const Entry& e, int q, double val)
{
    int i = e.i, j = e.j, k = e.k;
    switch(e.side)
    {
        case 1: i -= q + 1; break;
        case 4: i += q + 1; break;
        case 2: j += q + 1; break;
        case 3: j -= q + 1; break;
        case 5: k -= q + 1; break;
        case 6: k += q + 1; break;
    }
    f[d->id(i, j, k)] = val;
}

// dirichlet_ortho (gc_dirichlet-ortho.cpp:28-137), orderdir=2 (B23=1), wall value 0, no
// image point (needs dist < B29*dx, never true for the grids used here)
void dirichlet_ortho(orc_domain* d, double* f, const Entry& e)
{
    double dx;
    if(e.side == 1 || e.side == 4) dx = d->DP[0][e.i + MARGE];
    else if(e.side == 2 || e.side == 3) dx = d->DP[1][e.j + MARGE];
    else dx = d->DP[2][e.k + MARGE];
    int ys = 1;
    if(e.dist > d->DXM * (1.0 - 1.0e-6) && e.dist < d->DXM * (1.0 + 1.0e-6)) ys = 0;
    const int orderdir = 2;
    double pos[2] = {0.0, e.dist};
    double x[3];
    for(int m = 0; m < 3; ++m) x[m] = dx * double(m + 2 - ys);
    double y[8];
    y[0] = f[d->id(e.i, e.j, e.k)];
    y[1] = 0.0;
    for(int q = 0; q < 3; ++q)
    {
        y[orderdir + q] = 0.0;
        for(int m = 0; m < orderdir; ++m)
        {
            double weight = 1.0;
            for(int n = 0; n < orderdir; ++n)
                if(m != n) weight *= (x[q] - pos[n]) / (pos[m] - pos[n] + 1.0e-20);
            y[orderdir + q] += weight * y[m];
        }
    }
    for(int q = 0; q < 3; ++q) setg(d, f, e, q, y[orderdir + q - 1 + ys]);
}

// ghostcell::start1..4 (gc_start.cpp:27-182) + dgcpol1..4 (gc_dgcpol4.cpp:27-51)
void ghost(orc_domain* d, double* f, int kind, int gcv)
{
    const std::vector<Entry>& g = d->gcb[kind];
    const std::vector<Entry>& g1 = d->gcb[1];
    if(kind != 2 || d->jdir == 1)
        for(size_t q = 0; q < g.size(); ++q)
        {
            const Entry& e = g[q];
            if(e.side <= 0) continue;
            if(!((g1[q].side != 2 && g1[q].side != 3) || d->jdir == 1)) continue;
            int lb = label(kind, gcv, e.group, e.side);
            if((e.side == 2 || e.side == 3) && d->jdir == 0 && (lb == 4 || lb == 5)) continue;
            if(lb == 4)
            {
                double v = f[d->id(e.i, e.j, e.k)];
                for(int n = 0; n < 3; ++n) setg(d, f, e, n, v);
            }
            else if(lb == 5)
            {
                for(int n = 0; n < 3; ++n) setg(d, f, e, n, 0.0);
            }
            else if(lb == 1)
                dirichlet_ortho(d, f, e);
        }
    for(const Dgc& c : d->dgc[kind])
        f[d->id(c.i + c.di, c.j + c.dj, c.k + c.dk)] = f[d->id(c.i, c.j, c.k)];
}

// ------------------------------------------------------------------ WENO kernels
struct Weno
{
    const orc_domain* d;
    double q1, q2, q3, q4, q5;

    // is_min/max + weight_min/max + grad (weno_nug_func.h:48-196; weno_hj_df_nug.cpp:110-177)
    double recon(int axis, int ip, int uf, bool minus) const
    {
        const double* I = &d->isf[axis][(size_t)ip * 36 + uf * 18];
        const double* C = &d->cf[axis][(size_t)ip * 12 + uf * 6];
        const double* Q = &d->qf[axis][(size_t)ip * 24 + uf * 12];
        const double dq12 = q1 - q2, dq23 = q2 - q3, dq32 = q3 - q2, dq34 = q3 - q4, dq43 = q4 - q3,
                     dq54 = q5 - q4;
        int r = minus ? 0 : 3;
        double is1 = I[(r + 0) * 3 + 0] * dq54 * dq54 + I[(r + 0) * 3 + 1] * (dq54) * (dq34) + I[(r + 0) * 3 + 2] * dq34 * dq34;
        double is2 = I[(r + 1) * 3 + 0] * dq23 * dq23 + I[(r + 1) * 3 + 1] * (dq43) * (dq23) + I[(r + 1) * 3 + 2] * dq43 * dq43;
        double is3 = I[(r + 2) * 3 + 0] * dq12 * dq12 + I[(r + 2) * 3 + 1] * (dq32) * (dq12) + I[(r + 2) * 3 + 2] * dq32 * dq32;
        const double psi = 1.0e-6, eps = 0.0;
        const double a1 = is1 + psi, a2 = is2 + psi, a3 = is3 + psi;
        const double c0 = C[r], c1 = C[r + 1], c2 = C[r + 2];
        double w1 = c0 / (eps + (a1 * a1) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
        double w2 = c1 / (eps + (a2 * a2) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
        double w3 = c2 / (eps + (a3 * a3) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
        if(minus)
            return w1 * (q4 + Q[0] * (q3 - q4) - Q[1] * (q5 - q4))
                 + w2 * (q3 + Q[2] * (q4 - q3) - Q[3] * (q2 - q3))
                 + w3 * (q2 + Q[4] * (q1 - q2) + Q[5] * (q3 - q2));
        return w1 * (q4 + Q[6] * (q3 - q4) + Q[7] * (q5 - q4))
             + w2 * (q3 + Q[8] * (q2 - q3) - Q[9] * (q4 - q3))
             + w3 * (q2 + Q[10] * (q3 - q2) - Q[11] * (q1 - q2));
    }
};

inline long stride(const orc_domain* d, int axis) { return axis == 0 ? 1 : axis == 1 ? d->sx : d->sx * d->sy; }

// HJ-WENO gradient along one axis (weno_hj_df_nug.cpp:110-177 + iqmin_0/iqmax_0 :344-396;
// identical arithmetic in ddweno_nug_sf.cpp:39-224): divided differences over DXP.
double hj_grad(const orc_domain* d, const double* f, int i, int j, int k, int axis, double advec)
{
    if(!(advec > 0.0) && !(advec < 0.0)) return 0.0;
    Weno W{d};
    long s = stride(d, axis);
    long c = d->id(i, j, k);
    int ix = axis == 0 ? i : axis == 1 ? j : k;
    const double* D = d->DP[axis].data();
    int ip = ix + MARGE;
    if(advec > 0.0)
    {
        W.q1 = (f[c - 2 * s] - f[c - 3 * s]) / D[ip - 3];
        W.q2 = (f[c - s] - f[c - 2 * s]) / D[ip - 2];
        W.q3 = (f[c] - f[c - s]) / D[ip - 1];
        W.q4 = (f[c + s] - f[c]) / D[ip];
        W.q5 = (f[c + 2 * s] - f[c + s]) / D[ip + 1];
        return W.recon(axis, ip, 0, true);
    }
    W.q1 = (f[c - s] - f[c - 2 * s]) / D[ip - 2];
    W.q2 = (f[c] - f[c - s]) / D[ip - 1];
    W.q3 = (f[c + s] - f[c]) / D[ip];
    W.q4 = (f[c + 2 * s] - f[c + s]) / D[ip + 1];
    W.q5 = (f[c + 3 * s] - f[c + 2 * s]) / D[ip + 2];
    return W.recon(axis, ip, 0, false);
}

// weno_flux_nug::fx/fy/fz (weno_flux_nug.cpp:136-258) with iqmin..kqmax (:260-313):
// values, not differences; x upwind-max stencil repeats f(i+2) (weno_flux_nug.cpp:288-295).
double flux_face(const orc_domain* d, const double* f, int i, int j, int k, int axis, int uf, double advec)
{
    if(!(advec > 0.0) && !(advec < 0.0)) return 0.0;
    Weno W{d};
    long s = stride(d, axis);
    long c = d->id(i, j, k);
    int ip = (axis == 0 ? i : axis == 1 ? j : k) + MARGE;
    if(advec > 0.0)
    {
        W.q1 = f[c - 2 * s]; W.q2 = f[c - s]; W.q3 = f[c]; W.q4 = f[c + s]; W.q5 = f[c + 2 * s];
        return W.recon(axis, ip, uf, true);
    }
    W.q1 = f[c - s]; W.q2 = f[c]; W.q3 = f[c + s]; W.q4 = f[c + 2 * s];
    W.q5 = axis == 0 ? f[c + 2 * s] : f[c + 3 * s];
    return W.recon(axis, ip, uf, false);
}

// weno_hj_df_nug::start(ipol=4)/aij (weno_hj_df_nug.cpp:58-108) with
// flux_HJ_CDS2 ipol=4 cell-centred velocities (flux_HJ_CDS2.cpp:59-63,90-95,122-127)
void hj(orc_domain* d, const double* phi, const double* u, const double* v, const double* w, double* L)
{
    #pragma omp parallel for schedule(static) num_threads(d->threads) if(d->threads > 1)
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        double iadvec = 0.5 * (u[c] + u[c - 1]);
        double jadvec = 0.5 * (v[c] + v[c - d->sx]);
        double kadvec = 0.5 * (w[c] + w[c - d->sx * d->sy]);
        double Lc = -iadvec * hj_grad(d, phi, i, j, k, 0, iadvec);
        if(d->jdir == 1) Lc -= jadvec * hj_grad(d, phi, i, j, k, 1, jadvec);
        Lc -= kadvec * hj_grad(d, phi, i, j, k, 2, kadvec);
        L[c] += Lc;
    }
}

// weno_flux_nug::start/aij (weno_flux_nug.cpp:84-134) with flux_face_CDS2 face velocities
// (flux_face_CDS2.cpp:36-116)
void flux(orc_domain* d, int ipol, const double* q, const double* u, const double* v, const double* w, double* acc)
{
    const long sx = d->sx, sxy = d->sx * d->sy;
    int il = ipol == 1 ? d->last[1] : 0, jl = ipol == 2 ? d->last[2] : 0, kl = ipol == 3 ? d->last[3] : 0;
    const std::vector<int>& fl = d->flag[ipol];
    const double* DX = ipol == 1 ? d->DP[0].data() : d->DN[0].data();
    const double* DY = ipol == 2 ? d->DP[1].data() : d->DN[1].data();
    const double* DZ = ipol == 3 ? d->DP[2].data() : d->DN[2].data();
    int ufx = ipol == 1, ufy = ipol == 2, ufz = ipol == 3;
    #pragma omp parallel for schedule(static) num_threads(d->threads) if(d->threads > 1)
    for(int i = 0; i < d->nx - il; ++i)
    for(int j = 0; j < d->ny - jl; ++j)
    for(int k = 0; k < d->nz - kl; ++k)
    {
        long c = d->id(i, j, k);
        if(fl[c] <= 0) continue;
        double iv1, iv2, jv1, jv2, kv1, kv2;
        if(ipol == 1)
        {
            iv1 = 0.5 * (u[c] + u[c - 1]);               iv2 = 0.5 * (u[c] + u[c + 1]);
            jv1 = 0.5 * (v[c - sx] + v[c + 1 - sx]);      jv2 = 0.5 * (v[c] + v[c + 1]);
            kv1 = 0.5 * (w[c - sxy] + w[c + 1 - sxy]);    kv2 = 0.5 * (w[c] + w[c + 1]);
        }
        else if(ipol == 2)
        {
            iv1 = 0.5 * (u[c - 1] + u[c - 1 + sx]);       iv2 = 0.5 * (u[c] + u[c + sx]);
            jv1 = 0.5 * (v[c] + v[c - sx]);               jv2 = 0.5 * (v[c] + v[c + sx]);
            kv1 = 0.5 * (w[c - sxy] + w[c - sxy + sx]);   kv2 = 0.5 * (w[c] + w[c + sx]);
        }
        else
        {
            iv1 = 0.5 * (u[c - 1] + u[c - 1 + sxy]);      iv2 = 0.5 * (u[c] + u[c + sxy]);
            jv1 = 0.5 * (v[c - sx] + v[c - sx + sxy]);    jv2 = 0.5 * (v[c] + v[c + sxy]);
            kv1 = 0.5 * (w[c] + w[c - sxy]);              kv2 = 0.5 * (w[c] + w[c + sxy]);
        }
        double fv1 = 0.0, fv2 = 0.0;
        double fu1 = flux_face(d, q, i - 1, j, k, 0, ufx, iv1);
        double fu2 = flux_face(d, q, i, j, k, 0, ufx, iv2);
        if(d->jdir == 1)
        {
            fv1 = flux_face(d, q, i, j - 1, k, 1, ufy, jv1);
            fv2 = flux_face(d, q, i, j, k, 1, ufy, jv2);
        }
        double fw1 = flux_face(d, q, i, j, k - 1, 2, ufz, kv1);
        double fw2 = flux_face(d, q, i, j, k, 2, ufz, kv2);
        double Lc = -((iv2 * fu2 - iv1 * fu1) / DX[i + MARGE])
                  - ((jv2 * fv2 - jv1 * fv1) / DY[j + MARGE])
                  - ((kv2 * fw2 - kv1 * fw1) / DZ[k + MARGE]);
        acc[c] += Lc;
    }
}

// ------------------------------------------------------------------ reinitialisation
// reinidisc_f::start/disc (reinidisc_f.cpp:37-127) with ddweno_nug_sf (modus 0), F49 = 1;
// reinidisc_fsf::start/disc (reinidisc_fsf.cpp:37-141) with ddweno_nug, F49 = 0: the same
// per-cell update (ddweno_nug and ddweno_nug_sf in modus 0 read the same differences,
// ddweno_nug.cpp:39-213, ddweno_nug_sf.cpp:41-277) restricted to cells whose 7-point
// stencil has one sign; every other cell gets L = 0 (reinidisc_fsf.cpp:68-73)
void reini_disc(orc_domain* d, const double* f, double* L)
{
    const long sxy = d->sx * d->sy;
    #pragma omp parallel for schedule(static) num_threads(d->threads) if(d->threads > 1)
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        if(d->F49 == 0)
        {
            L[c] = 0.0;
            const bool pos = f[c] >= 0.0 && f[c + 1] >= 0.0 && f[c - 1] >= 0.0 && f[c + d->sx] >= 0.0
                          && f[c - d->sx] >= 0.0 && f[c + sxy] >= 0.0 && f[c - sxy] >= 0.0;
            const bool neg = f[c] < 0.0 && f[c + 1] < 0.0 && f[c - 1] < 0.0 && f[c + d->sx] < 0.0
                          && f[c - d->sx] < 0.0 && f[c + sxy] < 0.0 && f[c - sxy] < 0.0;
            if(!(pos || neg)) continue;
        }
        double dx = 0.0, dy = 0.0, dz = 0.0;
        double lsv = f[c];
        double lsSig = lsv / sqrt(lsv * lsv);
        if(fabs(lsv) < 1.0e-8) lsSig = 1.0;
        const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
        double xmin = (lsv - f[c - 1]) / d->DP[0][ip - 1];
        double xplus = (f[c + 1] - lsv) / d->DP[0][ip];
        if(xmin * lsSig > 0.0 && xplus * lsSig > -xmin * lsSig) dx = hj_grad(d, f, i, j, k, 0, 1.0);
        if(xplus * lsSig < 0.0 && xmin * lsSig < -xplus * lsSig) dx = hj_grad(d, f, i, j, k, 0, -1.0);
        if(xplus * lsSig > 0.0 && xmin * lsSig < 0.0) dx = 0.0;
        if(d->jdir == 1)
        {
            double ymin = (lsv - f[c - d->sx]) / d->DP[1][jp - 1];
            double yplus = (f[c + d->sx] - lsv) / d->DP[1][jp];
            if(ymin * lsSig > 0.0 && yplus * lsSig > -ymin * lsSig) dy = hj_grad(d, f, i, j, k, 1, 1.0);
            if(yplus * lsSig < 0.0 && ymin * lsSig < -yplus * lsSig) dy = hj_grad(d, f, i, j, k, 1, -1.0);
            if(yplus * lsSig > 0.0 && ymin * lsSig < 0.0) dy = 0.0;
        }
        const long sz = d->sx * d->sy;
        double zmin = (lsv - f[c - sz]) / d->DP[2][kp - 1];
        double zplus = (f[c + sz] - lsv) / d->DP[2][kp];
        if(zmin * lsSig > 0.0 && zplus * lsSig > -zmin * lsSig) dz = hj_grad(d, f, i, j, k, 2, 1.0);
        if(zplus * lsSig < 0.0 && zmin * lsSig < -zplus * lsSig) dz = hj_grad(d, f, i, j, k, 2, -1.0);
        if(zplus * lsSig > 0.0 && zmin * lsSig < 0.0) dz = 0.0;
        double dnorm = sqrt(dx * dx + dy * dy + dz * dz);
        double deltax;
        if(d->jdir == 0) deltax = (1.0 / 2.0) * (d->DN[0][ip] + d->DN[2][kp]);
        else deltax = (1.0 / 3.0) * (d->DN[0][ip] + d->DN[1][jp] + d->DN[2][kp]);
        double sign = lsv / sqrt(lsv * lsv + dnorm * dnorm * deltax * deltax);
        if(sign != sign) sign = 1.0;
        L[c] = -(sign * dnorm - sign);
    }
}

// reini_RK3::start/time_preproc (reini_RK3.cpp:77-160): for count>0 gcv = gcval_phi (52 for
// F50 = 2); the initial reinitialisation at count==0 runs with gcv = gcval_iniphi (50) after a
// start4(f, gcv) of its own (reini_RK3.cpp:93-101), which the caller does
void reini(orc_domain* d, double* f, int iters, int gcv = 52)
{
    long n = d->sx * d->sy * d->sz;
    std::vector<double> L(n, 0.0), frk1(n, 0.0), frk2(n, 0.0), dtf(n, 0.0);
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        double a = d->DP[0][i + MARGE], b = d->DP[1][j + MARGE], c = d->DP[2][k + MARGE];
        double m = d->jdir == 0 ? (a < c ? a : c) : (((a < b ? a : b) < c) ? (a < b ? a : b) : c);
        dtf[d->id(i, j, k)] = d->prm.F43 * m;
    }
    auto interior = [&](auto fn) {
        #pragma omp parallel for schedule(static) num_threads(d->threads) if(d->threads > 1)
        for(int i = 0; i < d->nx; ++i)
        for(int j = 0; j < d->ny; ++j)
        for(int k = 0; k < d->nz; ++k)
            fn(d->id(i, j, k));
    };
    for(int q = 0; q < iters; ++q)
    {
        reini_disc(d, f, L.data());
        interior([&](long c) { frk1[c] = f[c] + dtf[c] * L[c]; });
        ghost(d, frk1.data(), 4, gcv);
        reini_disc(d, frk1.data(), L.data());
        interior([&](long c) { frk2[c] = 0.75 * f[c] + 0.25 * frk1[c] + 0.25 * dtf[c] * L[c]; });
        ghost(d, frk2.data(), 4, gcv);
        reini_disc(d, frk2.data(), L.data());
        interior([&](long c) { f[c] = (1.0 / 3.0) * f[c] + (2.0 / 3.0) * frk2[c] + (2.0 / 3.0) * dtf[c] * L[c]; });
        ghost(d, f, 4, gcv);
    }
}

// ------------------------------------------------------------------ fluid properties
// fluid_update_fsf::start (fluid_update_fsf.cpp:39-104)
void fluid_update(orc_domain* d, const double* phi, double* ro, double* visc, double* vol)
{
    const double psi = d->psi;
    double H = 0.0, v1 = 0.0, v2 = 0.0;
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        if(phi[c] > psi) H = 1.0;
        if(phi[c] < -psi) H = 0.0;
        if(fabs(phi[c]) <= psi) H = 0.5 * (1.0 + phi[c] / psi + (1.0 / PI_REF) * sin((PI_REF * phi[c]) / psi));
        ro[c] = d->prm.rho_water * H + d->prm.rho_air * (1.0 - H);
        visc[c] = d->prm.visc_water * H + d->prm.visc_air * (1.0 - H);
        double vol_c = d->DN[0][i + MARGE] * d->DN[1][j + MARGE] * d->DN[2][k + MARGE];
        v1 += vol_c * (H - (1.0 - 1.0));
        v2 += vol_c * (1.0 - H - (1.0 - 1.0));
    }
    ghost(d, ro, 4, 1);
    ghost(d, visc, 4, 1);
    if(vol) { vol[0] = v1; vol[1] = v2; }
}

// density_f::roface (density_f.cpp:35-51)
inline double roface(const orc_domain* d, const double* phi, long c, long off)
{
    double phival = 0.5 * (phi[c] + phi[c + off]);
    double H = 0.0;
    if(phival > d->psi) H = 1.0;
    if(phival < -d->psi) H = 0.0;
    if(fabs(phival) <= d->psi) H = 0.5 * (1.0 + phival / d->psi + (1.0 / PI_REF) * sin((PI_REF * phival) / d->psi));
    return d->prm.rho_water * H + d->prm.rho_air * (1.0 - H);
}

// ------------------------------------------------------------------ projection
// poisson_pcorr::start (poisson_pcorr.cpp:78-231), porosity 1 (CPOR = PORVAL = 1)
void poisson(orc_domain* d, const double* phi, const double* press, double* M, double* rhs)
{
    const long N = d->cellnum;
    double *Mp = M, *Mn = M + N, *Ms = M + 2 * N, *Mw = M + 3 * N, *Me = M + 4 * N, *Mt = M + 5 * N, *Mb = M + 6 * N;
    const long sx = d->sx, sxy = d->sx * d->sy;
    const double yd = d->ydir;
    long n = 0;
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
        const double *DXP = d->DP[0].data(), *DXN = d->DN[0].data(), *DYP = d->DP[1].data(),
                     *DYN = d->DN[1].data(), *DZP = d->DP[2].data(), *DZN = d->DN[2].data();
        Mp[n] = (1.0 * 1.0) / (roface(d, phi, c, 1) * DXP[ip] * DXN[ip])
              + (1.0 * 1.0) / (roface(d, phi, c, -1) * DXP[ip - 1] * DXN[ip])
              + (1.0 * 1.0) / (roface(d, phi, c, sx) * DYP[jp] * DYN[jp]) * yd
              + (1.0 * 1.0) / (roface(d, phi, c, -sx) * DYP[jp - 1] * DYN[jp]) * yd
              + (1.0 * 1.0) / (roface(d, phi, c, sxy) * DZP[kp] * DZN[kp])
              + (1.0 * 1.0) / (roface(d, phi, c, -sxy) * DZP[kp - 1] * DZN[kp]);
        Mn[n] = -(1.0 * 1.0) / (roface(d, phi, c, 1) * DXP[ip] * DXN[ip]);
        Ms[n] = -(1.0 * 1.0) / (roface(d, phi, c, -1) * DXP[ip - 1] * DXN[ip]);
        Mw[n] = -(1.0 * 1.0) / (roface(d, phi, c, sx) * DYP[jp] * DYN[jp]) * yd;
        Me[n] = -(1.0 * 1.0) / (roface(d, phi, c, -sx) * DYP[jp - 1] * DYN[jp]) * yd;
        Mt[n] = -(1.0 * 1.0) / (roface(d, phi, c, sxy) * DZP[kp] * DZN[kp]);
        Mb[n] = -(1.0 * 1.0) / (roface(d, phi, c, -sxy) * DZP[kp - 1] * DZN[kp]);
        ++n;
    }
    const std::vector<int>& f4 = d->flag[4];
    n = 0;
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        if(f4[c - 1] < 0) { rhs[n] -= Ms[n] * press[c - 1]; Ms[n] = 0.0; }
        if(f4[c + 1] < 0) { rhs[n] -= Mn[n] * press[c + 1]; Mn[n] = 0.0; }
        if(f4[c - sx] < 0 && d->jdir == 1) { rhs[n] -= Me[n] * press[c - sx]; Me[n] = 0.0; }
        if(f4[c + sx] < 0 && d->jdir == 1) { rhs[n] -= Mw[n] * press[c + sx]; Mw[n] = 0.0; }
        if(f4[c - sxy] < 0) { rhs[n] -= Mb[n] * press[c - sxy]; Mb[n] = 0.0; }
        if(f4[c + sxy] < 0) { rhs[n] -= Mt[n] * press[c + sxy]; Mt[n] = 0.0; }
        ++n;
    }
}

// bicgstab_ijk::solve and helpers (bicgstab_ijk.cpp:125-395), FLEXLOOP of var
// (bicgstab_ijk.cpp:55-85, looping.h:55-57,159): var 1..3 over flag1..3 with ulast / vlast /
// wlast, 4 and 5 over flag4
struct Bicg
{
    orc_domain* d;
    const double* M;
    long N;                       // entries of the vec (matrix_diag rows)
    int var = 5;
    std::vector<double> rhs, x, rj, r0, pj, vj, sj, tj, ph, sh, aii;

    template<class F> void loop(F fn)
    {
        const int il = var == 1 ? d->last[1] : 0, jl = var == 2 ? d->last[2] : 0, kl = var == 3 ? d->last[3] : 0;
        const std::vector<int>& fl = d->flag[var <= 3 ? var : 4];
        long n = 0;
        for(int i = 0; i < d->nx - il; ++i)
        for(int j = 0; j < d->ny - jl; ++j)
        for(int k = 0; k < d->nz - kl; ++k)
        {
            const long c = d->id(i, j, k);
            if(fl[c] <= 0) continue;
            fn(c, n);
            ++n;
        }
    }
    // vec index of the first FLEXLOOP cell of every i slab (for the slab-parallel loops)
    std::vector<long> n0;
    void slabs()
    {
        const int il = var == 1 ? d->last[1] : 0;
        n0.assign(d->nx - il + 1, 0);
        loop([&](long c, long n) {
            const int i = (int)(c % d->sx) - 3;
            if(n0[i + 1] <= n) n0[i + 1] = n + 1;
        });
        for(int i = 1; i <= d->nx - il; ++i)
            if(n0[i] < n0[i - 1]) n0[i] = n0[i - 1];
    }
    // the cells of loop(), i slabs in parallel when the domain runs more than one thread;
    // fn writes only the cell it is given
    template<class F> void ploop(F fn)
    {
        const int il = var == 1 ? d->last[1] : 0, jl = var == 2 ? d->last[2] : 0, kl = var == 3 ? d->last[3] : 0;
        const std::vector<int>& fl = d->flag[var <= 3 ? var : 4];
        #pragma omp parallel for schedule(static) num_threads(d->threads) if(d->threads > 1)
        for(int i = 0; i < d->nx - il; ++i)
        {
            long n = n0[i];
            for(int j = 0; j < d->ny - jl; ++j)
            for(int k = 0; k < d->nz - kl; ++k)
            {
                const long c = d->id(i, j, k);
                if(fl[c] <= 0) continue;
                fn(c, n);
                ++n;
            }
        }
    }
    // K sums over the cells of loop(): with one thread in the reference's serial loop order
    // (bit-exact); with more, per-slab partial sums added in slab order (deterministic for
    // a given thread count, a different rounding from the reference's)
    template<class A, int K, class F> void psum(A* out, F fn)
    {
        if(d->threads <= 1)
        {
            A acc[K];
            for(int q = 0; q < K; ++q) acc[q] = 0.0;
            loop([&](long c, long n) { fn(c, n, acc); });
            for(int q = 0; q < K; ++q) out[q] = acc[q];
            return;
        }
        const int il = var == 1 ? d->last[1] : 0, jl = var == 2 ? d->last[2] : 0, kl = var == 3 ? d->last[3] : 0;
        const std::vector<int>& fl = d->flag[var <= 3 ? var : 4];
        const int ni = d->nx - il;
        std::vector<A> part((size_t)ni * K, (A)0.0);
        #pragma omp parallel for schedule(static) num_threads(d->threads)
        for(int i = 0; i < ni; ++i)
        {
            A* acc = part.data() + (size_t)i * K;
            long n = n0[i];
            for(int j = 0; j < d->ny - jl; ++j)
            for(int k = 0; k < d->nz - kl; ++k)
            {
                const long c = d->id(i, j, k);
                if(fl[c] <= 0) continue;
                fn(c, n, acc);
                ++n;
            }
        }
        for(int q = 0; q < K; ++q) out[q] = 0.0;
        for(int i = 0; i < ni; ++i)
            for(int q = 0; q < K; ++q) out[q] += part[(size_t)i * K + q];
    }
    double Ax(const std::vector<double>& xx, long c, long n) const
    {
        const long sx = d->sx, sxy = d->sx * d->sy;
        if(d->jdir == 1)
            return M[n] * xx[c] + M[N + n] * xx[c + 1] + M[2 * N + n] * xx[c - 1] + M[3 * N + n] * xx[c + sx]
                 + M[4 * N + n] * xx[c - sx] + M[5 * N + n] * xx[c + sxy] + M[6 * N + n] * xx[c - sxy];
        // bicgstab_ijk_2D.cpp:308-320 (5-point)
        return M[n] * xx[c] + M[N + n] * xx[c + 1] + M[2 * N + n] * xx[c - 1] + M[5 * N + n] * xx[c + sxy]
             + M[6 * N + n] * xx[c - sxy];
    }
    // A: accumulator type of the dot products; double reproduces the reference, long
    // double (the "dot_ext" control) perturbs only their rounding, to measure how far the
    // solution moves when a different summation order is used (a device reduction)
    template<class A> double res_calc()
    {
        A resi[1];
        psum<A, 1>(resi, [&](long c, long n, A* a) {
            double y = rhs[c] - (Ax(x, c, n));
            a[0] += y * y;
        });
        return sqrt((double)resi[0]) / double(d->cellnum);
    }
    int solve(double stop, int maxiter, double* resout)
    {
        slabs();
        return d->dot_ext ? solve_t<long double>(stop, maxiter, resout) : solve_t<double>(stop, maxiter, resout);
    }
    template<class A> int solve_t(double stop, int maxiter, double* resout)
    {
        int solveriter = 0;
        double residual = 1.0e9;
        ploop([&](long c, long n) { aii[c] = -1.0 / (M[n] + 1.0e-19); });
        double r_j, norm_r0, r_j1 = 0.0;
    restart:
        {
            A acc[1];
            ploop([&](long c, long n) { rj[c] = rhs[c] - (Ax(x, c, n)); });
            psum<A, 1>(acc, [&](long c, long, A* a) { r0[c] = pj[c] = rj[c]; a[0] += rj[c] * r0[c]; });
            r_j = (double)acc[0];
        }
        norm_r0 = sqrt(r_j);
        // negative stop: relative to the initial residual (harness mapping of a relative
        // tolerance onto N44, not a reference feature)
        if(stop < 0.0) stop = -stop * sqrt(r_j) / double(d->cellnum);
        if((residual >= stop) && (solveriter < maxiter))
        {
            do
            {
                A s3[3];
                ploop([&](long c, long) { ph[c] = pj[c] * aii[c]; });
                ploop([&](long c, long n) { vj[c] = Ax(ph, c, n); });
                psum<A, 3>(s3, [&](long c, long, A* a) {
                    a[0] += vj[c] * r0[c];
                    a[1] += vj[c] * vj[c];
                    a[2] += rj[c] * rj[c];
                });
                const double sigma = (double)s3[0];
                const double norm_vj = sqrt((double)s3[1]);
                const double norm_rj = sqrt((double)s3[2]);
                double alpha = r_j / sigma;
                if(fabs(sigma) <= (1.0e-12 * (norm_vj * norm_r0)))
                {
                    residual = res_calc<A>();
                    ++solveriter;
                    goto restart;
                }
                if((fabs(alpha) * norm_vj / (norm_rj == 0 ? 1.0e-15 : norm_rj)) <= 0.08)
                {
                    residual = res_calc<A>();
                    ++solveriter;
                    goto restart;
                }
                A nsj[1];
                psum<A, 1>(nsj, [&](long c, long, A* a) {
                    sj[c] = rj[c] - alpha * vj[c];
                    a[0] += sj[c] * sj[c];
                });
                const double norm_sj = sqrt((double)nsj[0]);
                if(norm_sj > stop)
                {
                    ploop([&](long c, long) { sh[c] = sj[c] * aii[c]; });
                    ploop([&](long c, long n) { tj[c] = Ax(sh, c, n); });
                    A ww[2];
                    psum<A, 2>(ww, [&](long c, long, A* a) {
                        a[0] += tj[c] * sj[c];
                        a[1] += tj[c] * tj[c];
                    });
                    double w = (double)ww[0] / ((double)ww[1] == 0 ? 1.0e-15 : (double)ww[1]);
                    A rj1[1];
                    psum<A, 1>(rj1, [&](long c, long, A* a) {
                        x[c] += alpha * ph[c] + w * sh[c];
                        rj[c] = sj[c] - w * tj[c];
                        a[0] += rj[c] * r0[c];
                    });
                    r_j1 = (double)rj1[0];
                    double beta = alpha * r_j1 / (w * r_j == 0 ? 1.0e-15 : (w * r_j));
                    ploop([&](long c, long) { pj[c] = rj[c] + beta * (pj[c] - w * vj[c]); });
                }
                if(norm_sj <= stop)
                {
                    A rj1[1];
                    psum<A, 1>(rj1, [&](long c, long, A* a) {
                        x[c] += alpha * ph[c];
                        rj[c] = sj[c];
                        a[0] += rj[c] * r0[c];
                    });
                    r_j1 = (double)rj1[0];
                }
                r_j = r_j1;
                A res[1];
                psum<A, 1>(res, [&](long c, long, A* a) { a[0] += rj[c] * rj[c]; });
                residual = sqrt((double)res[0]) / double(d->cellnum);
                ++solveriter;
            } while((residual >= stop) && (solveriter < maxiter));
        }
        if(resout) *resout = residual;
        return solveriter;
    }
};

// entries of the vec of var: its FLEXLOOP cells
long veclen(orc_domain* d, int var)
{
    Bicg b;
    b.d = d;
    b.var = var;
    long n = 0;
    b.loop([&](long, long) { ++n; });
    return n;
}

int bicgstab(orc_domain* d, double* xf, const double* M, const double* rhsvec, double stop, int maxiter, double* res,
             int var = 5)
{
    long n = d->sx * d->sy * d->sz;
    Bicg b;
    b.d = d;
    b.M = M;
    b.var = var;
    b.N = veclen(d, var);
    for(auto* v : {&b.rhs, &b.x, &b.rj, &b.r0, &b.pj, &b.vj, &b.sj, &b.tj, &b.ph, &b.sh, &b.aii})
        v->assign(n, 0.0);
    // var 1..3: the persistent private x of the diffusion solver (var 4, 5 cover every
    // interior cell, so the pressure solver's stale cells are only its ghosts, always 0)
    if(var <= 3)
    {
        if(d->psolv_x.size() != (size_t)n) d->psolv_x.assign(n, 0.0);
        b.x.swap(d->psolv_x);
    }
    b.loop([&](long c, long k) { b.x[c] = xf[c]; b.rhs[c] = rhsvec[k]; });
    int it = b.solve(stop, maxiter, res);
    b.loop([&](long c, long) { xf[c] = b.x[c]; });
    if(var <= 3) b.x.swap(d->psolv_x);
    return it;
}

// res_calc (bicgstab_ijk.cpp:314-338) without the 1/N: r = rhs - A x per interior cell,
// written to rout (field layout) when given; returns ||r||_2
double residual(orc_domain* d, const double* xf, const double* M, const double* rhsvec, double* rout)
{
    long n = d->sx * d->sy * d->sz;
    Bicg b;
    b.d = d;
    b.M = M;
    b.N = d->cellnum;
    b.x.assign(n, 0.0);
    b.loop([&](long c, long) { b.x[c] = xf[c]; });
    double resi = 0.0;
    b.loop([&](long c, long k) {
        const double y = rhsvec[k] - (b.Ax(b.x, c, k));
        if(rout) rout[c] = y;
        resi += y * y;
    });
    return sqrt(resi);
}

// pjm_corr::start (pjm_corr.cpp:90-183) with poisson_pcorr + bicgstab_ijk
int pjm(orc_domain* d, const double* phi, double* u, double* v, double* w, double* press, double* pcorr, double alpha)
{
    const long n = d->sx * d->sy * d->sz;
    const long N = d->cellnum;
    const double dt = d->prm.dt;
    const long sx = d->sx, sxy = d->sx * d->sy;
    ghost(d, u, 1, 7);
    ghost(d, v, 2, 8);
    ghost(d, w, 3, 9);
    std::vector<double> rhs(N, 0.0), M(7 * N, 0.0);
    for(long c = 0; c < n; ++c) pcorr[c] = 0.0;
    ghost(d, pcorr, 4, 1);
    long m = 0;
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        rhs[m] = -(u[c] - u[c - 1]) / (alpha * dt * d->DN[0][i + MARGE])
               - (v[c] - v[c - sx]) / (alpha * dt * d->DN[1][j + MARGE]) * d->ydir
               - (w[c] - w[c - sxy]) / (alpha * dt * d->DN[2][k + MARGE]);
        ++m;
    }
    poisson(d, phi, pcorr, M.data(), rhs.data());
    double res;
    int it = bicgstab(d, pcorr, M.data(), rhs.data(), d->prm.N44, d->prm.N46, &res);
    ghost(d, pcorr, 4, 40);
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        press[c] += pcorr[c];
    }
    ghost(d, press, 4, 40);
    // ucorr / vcorr / wcorr (pjm_corr.cpp:123-149)
    for(int i = 0; i < d->nx - d->last[1]; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        if(d->flag[1][c] <= 0) continue;
        u[c] -= alpha * dt * 1.0 * 1.0 * ((pcorr[c + 1] - pcorr[c]) / (d->DP[0][i + MARGE] * roface(d, phi, c, 1)));
    }
    if(d->jdir == 1)
        for(int i = 0; i < d->nx; ++i)
        for(int j = 0; j < d->ny - d->last[2]; ++j)
        for(int k = 0; k < d->nz; ++k)
        {
            long c = d->id(i, j, k);
            if(d->flag[2][c] <= 0) continue;
            v[c] -= alpha * dt * 1.0 * 1.0 * ((pcorr[c + sx] - pcorr[c]) / (d->DP[1][j + MARGE] * roface(d, phi, c, sx)));
        }
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz - d->last[3]; ++k)
    {
        long c = d->id(i, j, k);
        if(d->flag[3][c] <= 0) continue;
        w[c] -= alpha * dt * 1.0 * 1.0 * ((pcorr[c + sxy] - pcorr[c]) / (d->DP[2][k + MARGE] * roface(d, phi, c, sxy)));
    }
    return it;
}

// ------------------------------------------------------------------ time steps
template<class F> void vel_loop(orc_domain* d, int comp, F fn)
{
    int il = comp == 1 ? d->last[1] : 0, jl = comp == 2 ? d->last[2] : 0, kl = comp == 3 ? d->last[3] : 0;
    const std::vector<int>& fl = d->flag[comp];
    for(int i = 0; i < d->nx - il; ++i)
    for(int j = 0; j < d->ny - jl; ++j)
    for(int k = 0; k < d->nz - kl; ++k)
    {
        long c = d->id(i, j, k);
        if(fl[c] > 0) fn(c, i, j, k);
    }
}

// vel_loop with i slabs in parallel when the domain runs threads > 1 (fn writes only cell c)
template<class F> void vel_loop_p(orc_domain* d, int comp, F fn)
{
    int il = comp == 1 ? d->last[1] : 0, jl = comp == 2 ? d->last[2] : 0, kl = comp == 3 ? d->last[3] : 0;
    const std::vector<int>& fl = d->flag[comp];
    #pragma omp parallel for schedule(static) num_threads(d->threads) if(d->threads > 1)
    for(int i = 0; i < d->nx - il; ++i)
    for(int j = 0; j < d->ny - jl; ++j)
    for(int k = 0; k < d->nz - kl; ++k)
    {
        long c = d->id(i, j, k);
        if(fl[c] > 0) fn(c, i, j, k);
    }
}

template<class F> void cell_loop(orc_domain* d, F fn)
{
    #pragma omp parallel for schedule(static) num_threads(d->threads) if(d->threads > 1)
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
        fn(d->id(i, j, k));
}

// ediff2::diff_u / diff_v / diff_w (ediff2.cpp:78-190): the explicit viscous term of
// component comp at cell c, added to F after the convection.  Laminar: eddyv = 0, so each
// visc + eddyv sum of the reference equals visc (adding +0.0 is exact for visc > 0).
static double ediff_term(const orc_domain* d, int comp, const double* u, const double* v, const double* w,
                         const double* nu, long c, int i, int j, int k)
{
    const long X = 1, Y = d->sx, Z = d->sx * d->sy;
    const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
    const double *DXN = d->DN[0].data(), *DYN = d->DN[1].data(), *DZN = d->DN[2].data();
    const double *DXP = d->DP[0].data(), *DYP = d->DP[1].data(), *DZP = d->DP[2].data();
    auto V = [&](long o) { return nu[c + o]; };
    if(comp == 1)
    {
        const double uc = u[c];
        const double ddy_p = (((V(0) + V(X)) + V(Y)) + V(X + Y)) * 0.25;
        const double ddy_m = (((V(-Y) + V(X - Y)) + V(0)) + V(X)) * 0.25;
        const double ddz_p = (((V(0) + V(X)) + V(Z)) + V(X + Z)) * 0.25;
        const double ddz_m = (((V(-Z) + V(X - Z)) + V(0)) + V(X)) * 0.25;
        return 2.0 * ((u[c + X] - uc) * (V(X) / DXN[ip + 1]) - (uc - u[c - X]) * (V(0) / DXN[ip])) / DXP[ip]
               + ((u[c + Y] - uc) * (ddy_p / DYP[jp]) - (uc - u[c - Y]) * (ddy_m / DYP[jp - 1])) / DYN[jp]
               + ((u[c + Z] - uc) * (ddz_p / DZP[kp]) - (uc - u[c - Z]) * (ddz_m / DZP[kp - 1])) / DZN[kp]
               + ((v[c + X] - v[c]) * (ddy_p / DXP[ip]) - (v[c + X - Y] - v[c - Y]) * (ddy_m / DXP[ip])) / DYN[jp]
               + ((w[c + X] - w[c]) * (ddz_p / DXP[ip]) - (w[c + X - Z] - w[c - Z]) * (ddz_m / DXP[ip])) / DZN[kp];
    }
    if(comp == 2)
    {
        const double vc = v[c];
        const double ddx_p = (((V(0) + V(Y)) + V(X)) + V(X + Y)) * 0.25;
        const double ddx_m = (((V(-X) + V(Y - X)) + V(0)) + V(Y)) * 0.25;
        const double ddz_p = (((V(0) + V(Y)) + V(Z)) + V(Y + Z)) * 0.25;
        const double ddz_m = (((V(-Z) + V(Y - Z)) + V(0)) + V(Y)) * 0.25;
        return ((v[c + X] - vc) * (ddx_p / DXP[ip]) - (vc - v[c - X]) * (ddx_m / DXP[ip - 1])) / DXN[ip]
               + 2.0 * ((v[c + Y] - vc) * (V(Y) / DYN[jp + 1]) - (vc - v[c - Y]) * (V(0) / DYN[jp])) / DYP[jp]
               + ((v[c + Z] - vc) * (ddz_p / DZP[kp]) - (vc - v[c - Z]) * (ddz_m / DZP[kp - 1])) / DZN[kp]
               + ((u[c + Y] - u[c]) * (ddx_p / DYP[jp]) - (u[c - X + Y] - u[c - X]) * (ddx_m / DYP[jp])) / DXN[ip]
               + ((w[c + Y] - w[c]) * (ddz_p / DYP[jp]) - (w[c + Y - Z] - w[c - Z]) * (ddz_m / DYP[jp])) / DZN[kp];
    }
    const double wc = w[c];
    const double ddx_p = (((V(0) + V(Z)) + V(X)) + V(X + Z)) * 0.25;
    const double ddx_m = (((V(-X) + V(Z - X)) + V(0)) + V(Z)) * 0.25;
    const double ddy_p = (((V(0) + V(Z)) + V(Y)) + V(Y + Z)) * 0.25;
    const double ddy_m = (((V(-Y) + V(Z - Y)) + V(0)) + V(Z)) * 0.25;
    return ((w[c + X] - wc) * (ddx_p / DXP[ip]) - (wc - w[c - X]) * (ddx_m / DXP[ip - 1])) / DXN[ip]
           + ((w[c + Y] - wc) * (ddy_p / DYP[jp]) - (wc - w[c - Y]) * (ddy_m / DYP[jp - 1])) / DYN[jp]
           + 2.0 * ((w[c + Z] - wc) * (V(Z) / DZN[kp + 1]) - (wc - w[c - Z]) * (V(0) / DZN[kp])) / DZP[kp]
           + ((u[c + Z] - u[c]) * (ddx_p / DZP[kp]) - (u[c - X + Z] - u[c - X]) * (ddx_m / DZP[kp])) / DXN[ip]
           + ((v[c + Z] - v[c]) * (ddy_p / DZP[kp]) - (v[c - Y + Z] - v[c - Y]) * (ddy_m / DZP[kp])) / DYN[jp];
}

// idiff2_FS::diff_u / diff_v / diff_w (idiff2_FSu.cpp:40-278, idiff2_FSv.cpp:29-265,
// idiff2_FSw.cpp:29-265) and idiff2_FS_2D::diff_u / diff_w (idiff2_FSu_2D.cpp:40-218,
// idiff2_FSw_2D.cpp:29-207): the implicit viscous operator of component comp over its
// staggered loop, in vec order, into M (7 rows) and rhs (which the caller zeroed, as irhs
// leaves it).  Laminar: eddyv = 0, so each visc + eddyv of the reference equals visc;
// no porosity: CPOR = 1; D22 = 1 (ini.cpp:393): a row whose neighbour lies outside the
// component's grid moves that coupling times the cell's own value to the right-hand side,
// but only once the direct-forcing flags are set (DF1..3 > 0 on the staggered loops after
// the first ghostcell::solid_forcing_flag_update, momentum_forcing.cpp:102; they are 0
// before, so the rows stay unfolded).  diff_w reads the fdm's u and v (au, av) at k+1 and
// the argument u, v elsewhere (idiff2_FSw.cpp:94-95).
void idiff_assemble(orc_domain* d, int comp, double alpha, const double* vin, const double* u, const double* v,
                    const double* w, const double* au, const double* av, const double* nu, double* M, double* rhs)
{
    const long N = veclen(d, comp);
    double *Mp = M, *Mn = M + N, *Ms = M + 2 * N, *Mw = M + 3 * N, *Me = M + 4 * N, *Mt = M + 5 * N, *Mb = M + 6 * N;
    const long X = 1, Y = d->sx, Z = d->sx * d->sy;
    const double *DXN = d->DN[0].data(), *DYN = d->DN[1].data(), *DZN = d->DN[2].data();
    const double *DXP = d->DP[0].data(), *DYP = d->DP[1].data(), *DZP = d->DP[2].data();
    const double dt = d->prm.dt;
    const double CPOR = 1.0;
    const bool two = d->jdir == 0;
    const int ie = d->nx - (comp == 1 ? d->last[1] : 0), je = d->ny - (comp == 2 ? d->last[2] : 0),
              ke = d->nz - (comp == 3 ? d->last[3] : 0);
    auto outside = [&](int i, int j, int k) { return i < 0 || i >= ie || j < 0 || j >= je || k < 0 || k >= ke; };
    const double* own = comp == 1 ? u : comp == 2 ? v : w;
    long n = 0;
    vel_loop(d, comp, [&](long c, int i, int j, int k) {
        const int IP = i + MARGE, JP = j + MARGE, KP = k + MARGE;
        auto V = [&](long o) { return nu[c + o]; };
        if(comp == 1)
        {
            const double ddy_p = 0.25 * (((V(0) + V(X)) + V(Y)) + V(X + Y));
            const double ddy_m = 0.25 * (((V(-Y) + V(X - Y)) + V(0)) + V(X));
            const double ddz_p = 0.25 * (((V(0) + V(X)) + V(Z)) + V(X + Z));
            const double ddz_m = 0.25 * (((V(-Z) + V(X - Z)) + V(0)) + V(X));
            if(!two)
            {
                Mp[n] = 2.0 * V(X) / (DXN[IP + 1] * DXP[IP]) + 2.0 * V(0) / (DXN[IP] * DXP[IP])
                      + ddy_p / (DYP[JP] * DYN[JP]) + ddy_m / (DYP[JP - 1] * DYN[JP]) + ddz_p / (DZP[KP] * DZN[KP])
                      + ddz_m / (DZP[KP - 1] * DZN[KP]) + CPOR / (alpha * dt);
                rhs[n] += ((v[c + X] - v[c]) * ddy_p - (v[c + X - Y] - v[c - Y]) * ddy_m) / (DXP[IP] * DYN[JP])
                        + ((w[c + X] - w[c]) * ddz_p - (w[c + X - Z] - w[c - Z]) * ddz_m) / (DXP[IP] * DZN[KP])
                        + (CPOR * vin[c]) / (alpha * dt);
                Me[n] = -ddy_m / (DYP[JP - 1] * DYN[JP]);
                Mw[n] = -ddy_p / (DYP[JP] * DYN[JP]);
            }
            else
            {
                Mp[n] = 2.0 * V(X) / (DXN[IP + 1] * DXP[IP]) + 2.0 * V(0) / (DXN[IP] * DXP[IP])
                      + ddz_p / (DZP[KP] * DZN[KP]) + ddz_m / (DZP[KP - 1] * DZN[KP]) + CPOR / (alpha * dt);
                rhs[n] += ((w[c + X] - w[c]) * ddz_p - (w[c + X - Z] - w[c - Z]) * ddz_m) / (DXP[IP] * DZN[KP])
                        + (CPOR * vin[c]) / (alpha * dt);
            }
            Ms[n] = -2.0 * V(0) / (DXN[IP] * DXP[IP]);
            Mn[n] = -2.0 * V(X) / (DXN[IP + 1] * DXP[IP]);
            Mb[n] = -ddz_m / (DZP[KP - 1] * DZN[KP]);
            Mt[n] = -ddz_p / (DZP[KP] * DZN[KP]);
        }
        else if(comp == 2)
        {
            const double ddx_p = 0.25 * (((V(0) + V(Y)) + V(X)) + V(X + Y));
            const double ddx_m = 0.25 * (((V(-X) + V(Y - X)) + V(0)) + V(Y));
            const double ddz_p = 0.25 * (((V(0) + V(Y)) + V(Z)) + V(Y + Z));
            const double ddz_m = 0.25 * (((V(-Z) + V(Y - Z)) + V(0)) + V(Y));
            Mp[n] = 2.0 * V(Y) / (DYN[JP + 1] * DYP[JP]) + 2.0 * V(0) / (DYN[JP] * DYP[JP])
                  + ddx_p / (DXP[IP] * DXN[IP]) + ddx_m / (DXP[IP - 1] * DXN[IP]) + ddz_p / (DZP[KP] * DZN[KP])
                  + ddz_m / (DZP[KP - 1] * DZN[KP]) + CPOR / (alpha * dt);
            rhs[n] += ((u[c + Y] - u[c]) * ddx_p - (u[c - X + Y] - u[c - X]) * ddx_m) / (DYP[JP] * DXN[IP])
                    + ((w[c + Y] - w[c]) * ddz_p - (w[c + Y - Z] - w[c - Z]) * ddz_m) / (DYP[JP] * DZN[KP])
                    + (CPOR * vin[c]) / (alpha * dt);
            Ms[n] = -ddx_m / (DXP[IP - 1] * DXN[IP]);
            Mn[n] = -ddx_p / (DXP[IP] * DXN[IP]);
            Me[n] = -2.0 * V(0) / (DYN[JP] * DYP[JP]);
            Mw[n] = -2.0 * V(Y) / (DYN[JP + 1] * DYP[JP]);
            Mb[n] = -ddz_m / (DZP[KP - 1] * DZN[KP]);
            Mt[n] = -ddz_p / (DZP[KP] * DZN[KP]);
        }
        else
        {
            const double ddx_p = 0.25 * (((V(0) + V(Z)) + V(X)) + V(X + Z));
            const double ddx_m = 0.25 * (((V(-X) + V(Z - X)) + V(0)) + V(Z));
            if(!two)
            {
                const double ddy_p = 0.25 * (((V(0) + V(Z)) + V(Y)) + V(Y + Z));
                const double ddy_m = 0.25 * (((V(-Y) + V(Z - Y)) + V(0)) + V(Z));
                Mp[n] = 2.0 * V(Z) / (DZN[KP + 1] * DZP[KP]) + 2.0 * V(0) / (DZN[KP] * DZP[KP])
                      + ddx_p / (DXP[IP] * DXN[IP]) + ddx_m / (DXP[IP - 1] * DXN[IP]) + ddy_p / (DYP[JP] * DYN[JP])
                      + ddy_m / (DYP[JP - 1] * DYN[JP]) + CPOR / (alpha * dt);
                rhs[n] += ((au[c + Z] - u[c]) * ddx_p - (u[c - X + Z] - u[c - X]) * ddx_m) / (DZP[KP] * DXN[IP])
                        + ((av[c + Z] - v[c]) * ddy_p - (v[c - Y + Z] - v[c - Y]) * ddy_m) / (DZP[KP] * DYN[JP])
                        + (CPOR * vin[c]) / (alpha * dt);
                Me[n] = -ddy_m / (DYP[JP - 1] * DYN[JP]);
                Mw[n] = -ddy_p / (DYP[JP] * DYN[JP]);
            }
            else
            {
                Mp[n] = 2.0 * V(Z) / (DZN[KP + 1] * DZP[KP]) + 2.0 * V(0) / (DZN[KP] * DZP[KP])
                      + ddx_p / (DXP[IP] * DXN[IP]) + ddx_m / (DXP[IP - 1] * DXN[IP]) + CPOR / (alpha * dt);
                rhs[n] += ((u[c + Z] - u[c]) * ddx_p - (u[c - X + Z] - u[c - X]) * ddx_m) / (DZP[KP] * DXN[IP])
                        + (CPOR * vin[c]) / (alpha * dt);
            }
            Ms[n] = -ddx_m / (DXP[IP - 1] * DXN[IP]);
            Mn[n] = -ddx_p / (DXP[IP] * DXN[IP]);
            Mb[n] = -2.0 * V(0) / (DZN[KP] * DZP[KP]);
            Mt[n] = -2.0 * V(Z) / (DZN[KP + 1] * DZP[KP]);
        }
        // D22 = 1 wall folding (idiff2_FSu.cpp:120-170 and the v, w, 2-D forms): s, n, e, w,
        // b, t in that order, with the cell's own value of the component
        if(d->df_set)
        {
            const double x = own[c];
            if(outside(i - 1, j, k)) { rhs[n] -= Ms[n] * x; Ms[n] = 0.0; }
            if(outside(i + 1, j, k)) { rhs[n] -= Mn[n] * x; Mn[n] = 0.0; }
            if(!two)
            {
                if(outside(i, j - 1, k)) { rhs[n] -= Me[n] * x; Me[n] = 0.0; }
                if(outside(i, j + 1, k)) { rhs[n] -= Mw[n] * x; Mw[n] = 0.0; }
            }
            if(outside(i, j, k - 1)) { rhs[n] -= Mb[n] * x; Mb[n] = 0.0; }
            if(outside(i, j, k + 1)) { rhs[n] -= Mt[n] * x; Mt[n] = 0.0; }
        }
        ++n;
    });
}

// the whole diff_u / diff_v / diff_w call: diff = vin on the component's loop, its ghosts
// (start1..3 with gcval 10, 11, 12), the operator, bicgstab_ijk (bicgstab_ijk_2D in 2-D)
// with var comp and stop N43, the ghosts again; returns the iterations
int idiff(orc_domain* d, int comp, double alpha, double* diff, const double* vin, const double* u, const double* v,
          const double* w, const double* au, const double* av, const double* nu)
{
    const int gcv = comp == 1 ? 10 : comp == 2 ? 11 : 12;
    vel_loop(d, comp, [&](long c, int, int, int) { diff[c] = vin[c]; });
    ghost(d, diff, comp, gcv);
    const long N = veclen(d, comp);
    std::vector<double> M(7 * N, 0.0), rhs(N, 0.0);
    idiff_assemble(d, comp, alpha, vin, u, v, w, au, av, nu, M.data(), rhs.data());
    double res;
    const int it = bicgstab(d, diff, M.data(), rhs.data(), d->N43, d->prm.N46, &res, comp);
    ghost(d, diff, comp, gcv);
    return it;
}

// ------------------------------------------------------------------ iowave (B98 = 2)
// iowave::intriangle (iowave_dist_ini.cpp:141-190)
int intriangle(double Ax, double Ay, double Bx, double By, double Cx, double Cy, double x0, double y0)
{
    double Px = x0, Py = y0, Pz = -10.0, Qx = x0, Qy = y0, Qz = +10.0;
    double PQx = Qx - Px, PQy = Qy - Py, PQz = Qz - Pz;
    double Mx = PQy * Pz - PQz * Py;
    double My = PQz * Px - PQx * Pz;
    double u = PQz * (Cx * By - Cy * Bx) + Mx * (Cx - Bx) + My * (Cy - By);
    double v = PQz * (Ax * Cy - Ay * Cx) + Mx * (Ax - Cx) + My * (Ay - Cy);
    double w = PQz * (Bx * Ay - By * Ax) + Mx * (Bx - Ax) + My * (By - Ay);
    int check = 1;
    if(u == 0.0 && v == 0.0 && w == 0.0) check = 0;
    // the reference's condition, && binding tighter than ||
    if((u >= 0.0 && v >= 0.0 && w >= 0.0) || ((u < 0.0 && v < 0.0 && w < 0.0) && check == 1)) return 1;
    return 0;
}

double wmin(double a, double b) { return a < b ? a : b; }
double wmax(double a, double b) { return a > b ? a : b; }

// distbeach_ini / distgen_ini (iowave_dist_ini.cpp:27-139): corners 1..4 of one zone
void zone_corners(const double* S, const double* E, double dist, double (*C)[2])
{
    double dx = E[0] - S[0], dy = E[1] - S[1];
    double n1x = -dy, n1y = dx, n2x = dy, n2y = -dx;
    double l1 = sqrt(pow(n1x, 2.0) + pow(n1y, 2.0));
    double l2 = sqrt(pow(n2x, 2.0) + pow(n2y, 2.0));
    l1 = fabs(l1) > 1.0e-20 ? l1 : 1.0e20;
    l2 = fabs(l2) > 1.0e-20 ? l2 : 1.0e20;
    C[0][0] = S[0] + (n1x / l1) * dist;
    C[0][1] = S[1] + (n1y / l1) * dist;
    C[1][0] = S[0] + (n2x / l2) * dist;
    C[1][1] = S[1] + (n2y / l2) * dist;
    C[2][0] = E[0] + (n1x / l1) * dist;
    C[2][1] = E[1] + (n1y / l1) * dist;
    C[3][0] = E[0] + (n2x / l2) * dist;
    C[3][1] = E[1] + (n2y / l2) * dist;
}

// distgen / distbeach (iowave_dist.cpp:130-190) at (x0, y0): the distance to the zone's line
// inside the zone, 1e20 outside
double zone_dist(const double (*C)[2], const double* S, const double* E, double x0, double y0)
{
    double dist = 1.0e20;
    int test1 = intriangle(C[0][0], C[0][1], C[2][0], C[2][1], C[1][0], C[1][1], x0, y0);
    int test2 = intriangle(C[2][0], C[2][1], C[3][0], C[3][1], C[1][0], C[1][1], x0, y0);
    if(test1 == 1 || test2 == 1)
    {
        double denom = sqrt(pow(E[1] - S[1], 2.0) + pow(E[0] - S[0], 2.0));
        denom = denom > 1.0e-20 ? denom : 1.0e20;
        dist = wmin(fabs((E[1] - S[1]) * x0 - (E[0] - S[0]) * y0 + E[0] * S[1] - E[1] * S[0]) / denom, dist);
    }
    return dist;
}

const double EE_REF = 2.71828182846;   // definitions.h:38

// rb1_ext (iowave_relax_functions.cpp:28-95), one generation zone
double rb1_ext(const orc_domain* d, double x0, double y0)
{
    const orc_domain::Wave& W = d->wave;
    double x = -1.0e20, r = 0.0;
    int t1 = intriangle(W.G[0][0], W.G[0][1], W.G[2][0], W.G[2][1], W.G[1][0], W.G[1][1], x0, y0);
    int t2 = intriangle(W.G[2][0], W.G[2][1], W.G[3][0], W.G[3][1], W.G[1][0], W.G[1][1], x0, y0);
    if(!(t1 == 1 || t2 == 1)) return 1.0;
    double denom = sqrt(pow(W.Ge[1] - W.Gs[1], 2.0) + pow(W.Ge[0] - W.Gs[0], 2.0));
    denom = denom > 1.0e-20 ? denom : 1.0e20;
    double xdist = wmin(fabs((W.Ge[1] - W.Gs[1]) * x0 - (W.Ge[0] - W.Gs[0]) * y0 + W.Ge[0] * W.Gs[1] - W.Ge[1] * W.Gs[0]) /
                            denom,
                        1.0e20);
    x = wmax(1.0 - xdist / W.w.B96_1, x);
    x = wmax(x, 0.0);
    r = 1.0 - (exp(pow(x, 3.5)) - 1.0) / (EE_REF - 1.0);
    return r;
}

// rb3_ext (iowave_relax_functions.cpp:143-200), one beach zone
double rb3_ext(const orc_domain* d, double x0, double y0)
{
    const orc_domain::Wave& W = d->wave;
    int t1 = intriangle(W.B[0][0], W.B[0][1], W.B[2][0], W.B[2][1], W.B[1][0], W.B[1][1], x0, y0);
    int t2 = intriangle(W.B[2][0], W.B[2][1], W.B[3][0], W.B[3][1], W.B[1][0], W.B[1][1], x0, y0);
    if(!(t1 == 1 || t2 == 1)) return 1.0;
    double denom = sqrt(pow(W.Be[1] - W.Bs[1], 2.0) + pow(W.Be[0] - W.Bs[0], 2.0));
    denom = denom > 1.0e-20 ? denom : 1.0e20;
    double x = wmin(fabs((W.Be[1] - W.Bs[1]) * x0 - (W.Be[0] - W.Bs[0]) * y0 + W.Be[0] * W.Bs[1] - W.Be[1] * W.Bs[0]) /
                        denom,
                    1.0e20);
    double dist2 = W.w.B96_2;
    x = (dist2 - fabs(x)) / (dist2 * W.dist2_fac);
    x = wmax(x, 0.0);
    double r = 0.0;
    r += 1.0 - (exp(pow(x, 3.5)) - 1.0) / (EE_REF - 1.0);
    r /= double(1);
    return r;
}

// iowave::xgen (iowave_dist.cpp:27-44) at (x1, y1)
double xgen(const orc_domain* d, double x1, double y1)
{
    const orc_domain::Wave& W = d->wave;
    return fabs(y1 - W.tan_alpha * x1 + W.tan_alpha * W.x0 - W.y0) / sqrt(pow(W.tan_alpha, 2.0) + 1.0);
}

// iowave::iowave (iowave.cpp:32-263) and wave_lib_parameters (wave_lib_parameters.cpp:28-160)
// for linear waves (B92 = 2) with B91 or B93, the default zones (B106 = B107 = B108 = 0)
void wave_setup(orc_domain* d, const orc_wave_desc& w)
{
    orc_domain::Wave& W = d->wave;
    W.on = 1;
    W.w = w;
    W.phimean = w.F60;
    W.wdt = W.phimean;
    W.pshift = w.B120 * (PI_REF / 180.0);
    if(w.B91 == 1)
    {
        W.wa = 0.5 * w.B91_1;
        double wL = w.B91_2;
        W.wL = wL;
        W.wk = (2.0 * PI_REF) / (wL > 1.0e-20 ? wL : 1.0e20);
        W.ww = sqrt(fabs(9.81 * W.wk * tanh(W.wk * (W.wdt))));
        double wf = W.ww / (2.0 * PI_REF);
        W.wT = 1.0 / wf;
    }
    if(w.B93 == 1)
    {
        W.wa = 0.5 * w.B93_1;
        W.wT = w.B93_2;
        double wL0 = (9.81 / (2.0 * PI_REF)) * W.wT * W.wT;
        double k0 = (2.0 * PI_REF) / wL0;
        double S0 = sqrt(k0 * W.wdt) * (1.0 + (k0 * W.wdt) / 6.0 + (k0 * k0 * W.wdt * W.wdt) / 30.0);
        double wL = wL0 * tanh(S0);
        for(int qn = 0; qn < 500; ++qn) wL = wL0 * tanh(2.0 * PI_REF * W.wdt / wL);
        W.wL = wL;
        double wf = 1.0 / W.wT;
        W.ww = wf * 2.0 * PI_REF;
        W.wk = (2.0 * PI_REF) / (wL > 1.0e-20 ? wL : 1.0e20);
    }
    W.singamma = sin((w.B105_1) * (PI_REF / 180.0));
    W.cosgamma = cos((w.B105_1) * (PI_REF / 180.0));
    W.psi = 0.6 * d->DXM;
    W.epsi = 3.0 * d->DXM;
    W.dist2_fac = w.B99 == 1 ? 2.0 : 1.0;
    const double xmin = d->X[0][MARGE], xmax = d->X[0][d->nx + MARGE];
    const double ymin = d->X[1][MARGE], ymax = d->X[1][d->ny + MARGE];
    W.Gs[0] = xmin;
    W.Gs[1] = ymin - 10.0 * d->DXM;
    W.Ge[0] = xmin;
    W.Ge[1] = ymax + 10.0 * d->DXM;
    W.Bs[0] = xmax;
    W.Bs[1] = ymin - 10.0 * d->DXM;
    W.Be[0] = xmax;
    W.Be[1] = ymax + 10.0 * d->DXM;
    zone_corners(W.Bs, W.Be, w.B96_2, W.B);
    zone_corners(W.Gs, W.Ge, w.B96_1, W.G);
    W.x0 = xmin;
    W.y0 = ymin;
    double alpha = (w.B105_1 + 90.0) * (PI_REF / 180.0);
    W.tan_alpha = tan(alpha);
    const long n = d->sx * d->sy * d->sz;
    W.eta.assign((size_t)d->nx * d->ny, 0.0);
    W.uval.assign(n, 0.0);
    W.vval.assign(n, 0.0);
    W.wval.assign(n, 0.0);
    W.lsval.assign(n, 0.0);
}

// the ramp (iowave_relax_functions.cpp:239-254)
double wave_ramp(const orc_domain* d)
{
    const orc_domain::Wave& W = d->wave;
    double f = 1.0;
    if(W.w.B101 == 1 && d->simtime < W.w.B102 * W.wT)
        f = d->simtime / (W.w.B102 * W.wT) - (1.0 / PI_REF) * sin(PI_REF * (d->simtime / (W.w.B102 * W.wT)));
    if(W.w.B101 == 2 && d->simtime < W.w.B102)
        f = d->simtime / (W.w.B102) - (1.0 / PI_REF) * sin(PI_REF * (d->simtime / (W.w.B102)));
    return f;
}

// wave_lib_linear (wave_lib_linear.cpp:47-85) through wave_interface (wave_interface.cpp:
// 137-219): wts = 0, wte = 1e20, wavetime = simtime
double wave_horzvel(const orc_domain* d, double x, double z)
{
    const orc_domain::Wave& W = d->wave;
    double teta = W.wk * x - W.ww * (d->simtime) + W.pshift;
    return W.ww * W.wa * (cosh(W.wk * (W.wdt + z)) / sinh(W.wk * W.wdt)) * cos(teta);
}
double wave_u(const orc_domain* d, double x, double z)
{
    z = wmax(z, -d->wave.phimean);
    double uvel = 0.0;
    if(d->simtime >= 0.0 && d->simtime <= 1.0e20) uvel = d->wave.cosgamma * wave_horzvel(d, x, z);
    return uvel;
}
double wave_v(const orc_domain* d, double x, double z)
{
    z = wmax(z, -d->wave.phimean);
    double vvel = 0.0;
    if(d->simtime >= 0.0 && d->simtime <= 1.0e20) vvel = d->wave.singamma * wave_horzvel(d, x, z);
    return vvel;
}
double wave_w(const orc_domain* d, double x, double z)
{
    const orc_domain::Wave& W = d->wave;
    z = wmax(z, -W.phimean);
    double wvel = 0.0;
    if(d->simtime >= 0.0 && d->simtime <= 1.0e20)
    {
        double teta = W.wk * x - W.ww * (d->simtime) + W.pshift;
        wvel = W.ww * W.wa * (sinh(W.wk * (W.wdt + z)) / sinh(W.wk * W.wdt)) * sin(teta);
    }
    return wvel;
}
double wave_eta(const orc_domain* d, double x)
{
    const orc_domain::Wave& W = d->wave;
    double eta = 0.0;
    if(d->simtime >= 0.0 && d->simtime <= 1.0e20) eta = W.wa * cos(W.wk * x - W.ww * (d->simtime) + W.pshift);
    return eta;
}

// iowave::wavegen_precalc_relax (iowave_precalc_relax.cpp:27-160) at d->simtime: eta in the
// generation zone's columns (gcsl_start4 fills only slice ghosts, which the relaxation loops
// do not read), then the target values of every generation-zone cell of the u, v, w and
// cell loops, kept per cell instead of by count
void wave_precalc(orc_domain* d)
{
    orc_domain::Wave& W = d->wave;
    const int nx = d->nx, ny = d->ny;
    const double *XN = d->X[0].data(), *YN = d->X[1].data(), *XP = d->P[0].data(), *YP = d->P[1].data();
    const double *ZN = d->X[2].data(), *ZP = d->P[2].data();
    auto dg = [&](int i, int j) { return zone_dist(W.G, W.Gs, W.Ge, XP[i + MARGE], YP[j + MARGE]); };
    auto eta = [&](int i, int j) -> double& { return W.eta[(size_t)i + (size_t)nx * j]; };
    for(int i = 0; i < nx; ++i)
    for(int j = 0; j < ny; ++j)
        if(W.w.B98 == 2 && dg(i, j) < 1.0e20) eta(i, j) = wave_eta(d, xgen(d, XP[i + MARGE], YP[j + MARGE]));
    vel_loop(d, 1, [&](long c, int i, int j, int k) {
        double xg = xgen(d, XN[i + 1 + MARGE], YP[j + MARGE]);
        double zloc1 = ZP[k + MARGE];
        double fsfloc = 0.5 * (eta(i, j) + eta(i + 1, j)) + W.phimean, z = 0.0;
        if(zloc1 <= fsfloc) z = zloc1 - W.phimean;
        if(zloc1 > fsfloc) z = 0.5 * (eta(i, j) + eta(i + 1, j));
        if(W.w.B98 == 2 && dg(i, j) < 1.0e20)
        {
            if(zloc1 <= fsfloc + W.epsi) W.uval[c] = wave_u(d, xg, z) + 0.0;
            if(zloc1 > fsfloc + W.epsi) W.uval[c] = 0.0;
        }
    });
    vel_loop(d, 2, [&](long c, int i, int j, int k) {
        double xg = xgen(d, XP[i + MARGE], YN[j + 1 + MARGE]);
        double zloc2 = ZP[k + MARGE];
        double fsfloc = 0.5 * (eta(i, j) + eta(i, j + 1)) + W.phimean, z = 0.0;
        if(zloc2 <= fsfloc) z = zloc2 - W.phimean;
        if(zloc2 > fsfloc) z = 0.5 * (eta(i, j) + eta(i, j + 1));
        if(W.w.B98 == 2 && dg(i, j) < 1.0e20)
        {
            if(zloc2 <= fsfloc + W.epsi) W.vval[c] = wave_v(d, xg, z);
            if(zloc2 > fsfloc + W.epsi) W.vval[c] = 0.0;
        }
    });
    vel_loop(d, 3, [&](long c, int i, int j, int k) {
        double xg = xgen(d, XP[i + MARGE], YP[j + MARGE]);
        double zloc3 = ZN[k + 1 + MARGE];
        double fsfloc = eta(i, j) + W.phimean, z = 0.0;
        if(zloc3 <= fsfloc) z = zloc3 - W.phimean;
        if(zloc3 > fsfloc) z = eta(i, j);
        if(W.w.B98 == 2 && dg(i, j) < 1.0e20)
        {
            if(zloc3 <= fsfloc + W.epsi) W.wval[c] = wave_w(d, xg, z);
            if(zloc3 > fsfloc + W.epsi) W.wval[c] = 0.0;
        }
    });
    for(int i = 0; i < nx; ++i)
    for(int j = 0; j < ny; ++j)
    for(int k = 0; k < d->nz; ++k)
        if(W.w.B98 == 2 && dg(i, j) < 1.0e20) W.lsval[d->id(i, j, k)] = eta(i, j) + W.phimean - ZP[k + MARGE];
}

// the band value H (= G) of the relaxation of velocity (iowave_relax.cpp:38-55): 1 above
// -psi, 0 below -epsi, a smoothed step between
void wave_HG(const orc_domain* d, double phival, double& H, double& G)
{
    const orc_domain::Wave& W = d->wave;
    H = G = 0.0;
    if(phival >= -W.psi) { H = 1.0; G = 1.0; }
    if(phival < -W.epsi) { H = 0.0; G = 0.0; }
    if(phival >= -W.epsi && phival < -W.psi)
    {
        H = 0.5 * (1.0 + phival / W.epsi + (1.0 / PI_REF) * sin((PI_REF * phival) / W.epsi));
        G = H;
    }
}

// iowave::u_relax / v_relax / w_relax (iowave_relax.cpp:28-330): comp 1..3 of field vel,
// with the band of phi on the component's face
void wave_vel_relax(orc_domain* d, int comp, const double* phi, double* vel)
{
    orc_domain::Wave& W = d->wave;
    const double *XP = d->P[0].data(), *YP = d->P[1].data(), *XN = d->X[0].data(), *YN = d->X[1].data();
    const long off = comp == 1 ? 1 : comp == 2 ? d->sx : d->sx * d->sy;
    const double* val = comp == 1 ? W.uval.data() : comp == 2 ? W.vval.data() : W.wval.data();
    const double ramp = wave_ramp(d);
    const bool beach = W.w.B99 == 1 || W.w.B99 == 2;
    vel_loop(d, comp, [&](long c, int i, int j, int k) {
        double x0 = XP[i + MARGE], y0 = YP[j + MARGE];
        double dg = zone_dist(W.G, W.Gs, W.Ge, x0, y0);
        double db = zone_dist(W.B, W.Bs, W.Be, x0, y0);
        double phival = 0.5 * (phi[c] + phi[c + off]);
        double H, G;
        wave_HG(d, phival, H, G);
        // relax1/2/4 on the component's point (pos1, pos2, pos3 = cell centre in x, y)
        double xr = comp == 1 ? XN[i + 1 + MARGE] : XP[i + MARGE];
        double yr = comp == 2 ? YN[j + 1 + MARGE] : YP[j + MARGE];
        if(W.w.B98 == 2 && dg < 1.0e20)
        {
            double r = rb1_ext(d, xr, yr);
            vel[c] = (1.0 - r) * ramp * val[c] * H + r * H * vel[c] + (1.0 - G) * vel[c];
            if(comp == 1) vel[c] += (1.0 - r) * ramp * W.w.W50 * (1.0 - H) + r * (1.0 - H) * W.w.W50;
        }
        if(beach && db < (comp == 1 ? 1.0e19 : 1.0e20)) vel[c] = rb3_ext(d, xr, yr) * vel[c];
    });
}

// iowave::p_relax (iowave_relax.cpp:332-356)
void wave_p_relax(orc_domain* d, const double* ro, double* press)
{
    orc_domain::Wave& W = d->wave;
    if(!(W.w.B99 == 1 || W.w.B99 == 2)) return;
    const double *XP = d->P[0].data(), *YP = d->P[1].data(), *ZP = d->P[2].data();
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        double x0 = XP[i + MARGE], y0 = YP[j + MARGE];
        if(zone_dist(W.B, W.Bs, W.Be, x0, y0) < 1.0e20)
        {
            double r = rb3_ext(d, x0, y0);
            press[c] = (1.0 - r) * ((W.phimean - ZP[k + MARGE]) * ro[c] * fabs(d->prm.gk)) + r * press[c];
        }
    }
}

// iowave::phi_relax (iowave_relax.cpp:358-395)
void wave_phi_relax(orc_domain* d, double* f)
{
    orc_domain::Wave& W = d->wave;
    const double *XP = d->P[0].data(), *YP = d->P[1].data(), *ZP = d->P[2].data();
    const double ramp = wave_ramp(d);
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        long c = d->id(i, j, k);
        double x0 = XP[i + MARGE], y0 = YP[j + MARGE];
        if(W.w.B98 == 2 && zone_dist(W.G, W.Gs, W.Ge, x0, y0) < 1.0e20)
        {
            double r = rb1_ext(d, x0, y0);
            f[c] = (1.0 - r) * ramp * W.lsval[c] + r * f[c];
        }
        if(W.w.B99 == 2 && zone_dist(W.B, W.Bs, W.Be, x0, y0) < 1.0e20)
        {
            double r = rb3_ext(d, x0, y0);
            f[c] = (1.0 - r) * (W.phimean - ZP[k + MARGE]) + r * f[c];
        }
    }
}

// u_relax, v_relax, w_relax and p_relax in the order momentum_FC3 calls them
// (momentum_FC3.cpp:218-221)
void wave_relax_uvwp(orc_domain* d, const double* phi, const double* ro, double* u, double* v, double* w,
                     double* press)
{
    wave_vel_relax(d, 1, phi, u);
    wave_vel_relax(d, 2, phi, v);
    wave_vel_relax(d, 3, phi, w);
    if(press) wave_p_relax(d, ro, press);
}

// momentum_FC3::start (momentum_FC3.cpp:122-430) with D20=0 (diff_void) or 1 (ediff2), B10=0,
// laminar (kepsilon_void zeroes F/G/H), void ioflow and forcing, pjm_corr projection.
// The working field ls is a member of momentum_FC3 whose ghost cells are never filled,
// so the stage-1 HJ-WENO reads zeros there (momentum_FC3.cpp:132-139).
struct FC3
{
    orc_domain* d;
    long n;
    std::vector<double> ls, frk1, frk2, L, F, urk1, urk2, vrk1, vrk2, wrk1, wrk2, udiff, vdiff, wdiff, pcorr;
    const double* nu = nullptr;     // the mixture viscosity (fluid_update) the diffusion reads

    explicit FC3(orc_domain* dd) : d(dd), n(dd->sx * dd->sy * dd->sz)
    {
        for(auto* v : {&ls, &frk1, &frk2, &L, &F, &urk1, &urk2, &vrk1, &vrk2, &wrk1, &wrk2, &udiff, &vdiff, &wdiff, &pcorr})
            v->assign(n, 0.0);
    }

    // F/G/H for component comp of velocity set (uu,vv,ww), pressure gradient and gravity
    void rhs_comp(int comp, const double* phi, const double* press, const double* uu, const double* vv, const double* ww)
    {
        const double* q = comp == 1 ? uu : comp == 2 ? vv : ww;
        long off = comp == 1 ? 1 : comp == 2 ? d->sx : d->sx * d->sy;
        const double* D = d->DP[comp - 1].data();
        double g = comp == 1 ? d->prm.gi : comp == 2 ? d->prm.gj : d->prm.gk;
        vel_loop_p(d, comp, [&](long c, int i, int j, int k) { F[c] = 0.0; });
        // pjm_corr::upgrad/vpgrad/wpgrad (pjm_corr.cpp:185-204)
        vel_loop_p(d, comp, [&](long c, int i, int j, int k) {
            int ix = comp == 1 ? i : comp == 2 ? j : k;
            F[c] -= 1.0 * (press[c + off] - press[c]) / (D[ix + MARGE] * roface(d, phi, c, off));
        });
        // irhs/jrhs/krhs (momentum_FC3.cpp:432-472), rhsvec = 0, Fext = 0, W29 = 0
        vel_loop_p(d, comp, [&](long c, int, int, int) { F[c] += (0.0 + g + 0.0 + 0.0) * 1.0; });
        flux(d, comp, q, uu, vv, ww, F.data());
        // pdiff->diff_u / diff_v / diff_w (momentum_FC3.cpp:170, 187, 204 and the stages)
        if(d->D20 == 1)
            vel_loop_p(d, comp, [&](long c, int i, int j, int k) { F[c] += ediff_term(d, comp, uu, vv, ww, nu, c, i, j, k); });
    }

    int step(double* phi, double* u, double* v, double* w, double* press, double* ro, double* visc, double* vol)
    {
        const double dt = d->prm.dt;
        int iters = 0;
        nu = visc;
        // ---- stage 1
        cell_loop(d, [&](long c) { L[c] = 0.0; ls[c] = phi[c]; });
        hj(d, ls.data(), u, v, w, L.data());
        cell_loop(d, [&](long c) { frk1[c] = ls[c] + dt * L[c]; });
        if(d->wave.on) wave_phi_relax(d, frk1.data());      // pflow->phi_relax (momentum_FC3.cpp:145)
        ghost(d, frk1.data(), 4, 52);
        reini(d, frk1.data(), 2);
        cell_loop(d, [&](long c) { phi[c] = frk1[c]; });
        ghost(d, phi, 4, 52);
        fluid_update(d, phi, ro, visc, vol);
        velocity_stage(phi, press, u, v, w, u, v, w, urk1, vrk1, wrk1, 1.0, 0);
        // momentum_forcing_start (momentum_FC3.cpp:212) ends with solid_forcing_flag_update:
        // the direct-forcing flags the implicit diffusion reads are set from here on
        d->df_set = 1;
        iters += pjm(d, phi, urk1.data(), vrk1.data(), wrk1.data(), press, pcorr.data(), 1.0);
        if(d->wave.on) wave_relax_uvwp(d, phi, ro, urk1.data(), vrk1.data(), wrk1.data(), press);   // :218-221
        ghost(d, urk1.data(), 1, 10);
        ghost(d, vrk1.data(), 2, 11);
        ghost(d, wrk1.data(), 3, 12);
        fluid_update(d, phi, ro, visc, vol);
        // ---- stage 2
        cell_loop(d, [&](long c) { L[c] = 0.0; });
        hj(d, frk1.data(), urk1.data(), vrk1.data(), wrk1.data(), L.data());
        cell_loop(d, [&](long c) { frk2[c] = 0.75 * ls[c] + 0.25 * frk1[c] + 0.25 * dt * L[c]; });
        if(d->wave.on) wave_phi_relax(d, frk2.data());      // momentum_FC3.cpp:244
        ghost(d, frk2.data(), 4, 52);
        reini(d, frk2.data(), 2);
        cell_loop(d, [&](long c) { phi[c] = frk2[c]; });
        ghost(d, phi, 4, 52);
        reini(d, phi, 2);
        fluid_update(d, phi, ro, visc, vol);
        velocity_stage(phi, press, u, v, w, urk1.data(), vrk1.data(), wrk1.data(), urk2, vrk2, wrk2, 0.25, 1);
        iters += pjm(d, phi, urk2.data(), vrk2.data(), wrk2.data(), press, pcorr.data(), 0.25);
        if(d->wave.on) wave_relax_uvwp(d, phi, ro, urk2.data(), vrk2.data(), wrk2.data(), press);   // :320-323
        ghost(d, urk2.data(), 1, 10);
        ghost(d, vrk2.data(), 2, 11);
        ghost(d, wrk2.data(), 3, 12);
        fluid_update(d, phi, ro, visc, vol);
        // ---- stage 3
        cell_loop(d, [&](long c) { L[c] = 0.0; });
        hj(d, frk2.data(), urk2.data(), vrk2.data(), wrk2.data(), L.data());
        cell_loop(d, [&](long c) { ls[c] = (1.0 / 3.0) * ls[c] + (2.0 / 3.0) * frk2[c] + (2.0 / 3.0) * dt * L[c]; });
        if(d->wave.on) wave_phi_relax(d, ls.data());        // momentum_FC3.cpp:346
        cell_loop(d, [&](long c) { phi[c] = ls[c]; });
        ghost(d, phi, 4, 52);
        reini(d, phi, 3);
        ghost(d, phi, 4, 52);
        fluid_update(d, phi, ro, visc, vol);
        std::vector<double> un(u, u + n), vn(v, v + n), wn(w, w + n);
        velocity_stage3(phi, press, u, v, w);
        iters += pjm(d, phi, u, v, w, press, pcorr.data(), 2.0 / 3.0);
        if(d->wave.on) wave_relax_uvwp(d, phi, ro, u, v, w, press);    // momentum_FC3.cpp:418-421
        ghost(d, u, 1, 10);
        ghost(d, v, 2, 11);
        ghost(d, w, 3, 12);
        fluid_update(d, phi, ro, visc, vol);
        return iters;
    }

    // stage 1 (kind 0) and stage 2 (kind 1) velocity predictors
    void velocity_stage(const double* phi, const double* press, const double* un, const double* vn, const double* wn,
                        const double* uu, const double* vv, const double* ww, std::vector<double>& ur,
                        std::vector<double>& vr, std::vector<double>& wr, double alpha, int kind)
    {
        const double dt = d->prm.dt;
        std::vector<double>* out[4] = {nullptr, &ur, &vr, &wr};
        std::vector<double>* dif[4] = {nullptr, &udiff, &vdiff, &wdiff};
        const double* in[4] = {nullptr, uu, vv, ww};
        const double* old[4] = {nullptr, un, vn, wn};
        for(int comp = 1; comp <= 3; ++comp)
        {
            if(comp == 2 && d->jdir == 0)
            {
                // jsource zeroes G but VLOOP is empty in 2D; nothing else runs
                continue;
            }
            rhs_comp(comp, phi, press, uu, vv, ww);
            std::vector<double>& df = *dif[comp];
            if(d->D20 == 2)
                // pdiff->diff_u/v/w with idiff2_FS: the fdm's u, v are the step's (un, vn)
                idiff(d, comp, alpha, df.data(), in[comp], uu, vv, ww, un, vn, nu);
            else
            {
                vel_loop_p(d, comp, [&](long c, int, int, int) { df[c] = in[comp][c]; });
                ghost(d, df.data(), comp, comp == 1 ? 10 : comp == 2 ? 11 : 12);
            }
            std::vector<double>& o = *out[comp];
            if(kind == 0)
                vel_loop_p(d, comp, [&](long c, int, int, int) { o[c] = df[c] + dt * 1.0 * F[c]; });
            else
                vel_loop_p(d, comp, [&](long c, int, int, int) {
                    o[c] = 0.75 * old[comp][c] + 0.25 * df[c] + 0.25 * dt * 1.0 * F[c];
                });
        }
    }

    void velocity_stage3(const double* phi, const double* press, double* u, double* v, double* w)
    {
        const double dt = d->prm.dt;
        double* tgt[4] = {nullptr, u, v, w};
        std::vector<double>* dif[4] = {nullptr, &udiff, &vdiff, &wdiff};
        const double* in[4] = {nullptr, urk2.data(), vrk2.data(), wrk2.data()};
        for(int comp = 1; comp <= 3; ++comp)
        {
            if(comp == 2 && d->jdir == 0) continue;
            rhs_comp(comp, phi, press, urk2.data(), vrk2.data(), wrk2.data());
            std::vector<double>& df = *dif[comp];
            if(d->D20 == 2)
                // the fdm's u and v are already this stage's results when w is solved
                idiff(d, comp, 2.0 / 3.0, df.data(), in[comp], urk2.data(), vrk2.data(), wrk2.data(), u, v, nu);
            else
            {
                vel_loop_p(d, comp, [&](long c, int, int, int) { df[c] = in[comp][c]; });
                ghost(d, df.data(), comp, comp == 1 ? 10 : comp == 2 ? 11 : 12);
            }
            double* t = tgt[comp];
            vel_loop_p(d, comp, [&](long c, int, int, int) {
                t[c] = (1.0 / 3.0) * t[c] + (2.0 / 3.0) * df[c] + (2.0 / 3.0) * dt * 1.0 * F[c];
            });
        }
    }
};

// levelset_RK3::start (levelset_RK3.cpp:103-160), phi transport stages only
void phi_rk3(orc_domain* d, double* ls, const double* u, const double* v, const double* w)
{
    const long n = d->sx * d->sy * d->sz;
    const double dt = d->prm.dt;
    std::vector<double> L(n, 0.0), ark1(n, 0.0), ark2(n, 0.0);
    cell_loop(d, [&](long c) { L[c] = 0.0; });
    hj(d, ls, u, v, w, L.data());
    cell_loop(d, [&](long c) { ark1[c] = ls[c] + dt * L[c]; });
    ghost(d, ark1.data(), 4, 52);
    cell_loop(d, [&](long c) { L[c] = 0.0; });
    hj(d, ark1.data(), u, v, w, L.data());
    cell_loop(d, [&](long c) { ark2[c] = 0.75 * ls[c] + 0.25 * ark1[c] + 0.25 * dt * L[c]; });
    ghost(d, ark2.data(), 4, 52);
    cell_loop(d, [&](long c) { L[c] = 0.0; });
    hj(d, ark2.data(), u, v, w, L.data());
    cell_loop(d, [&](long c) { ls[c] = (1.0 / 3.0) * ls[c] + (2.0 / 3.0) * ark2[c] + (2.0 / 3.0) * dt * L[c]; });
    ghost(d, ls, 4, 52);
}

// one level-set benchmark step: driver_loop_cfd.cpp:82-96 with levelset_RK3 (F30=3),
// momentum_RK3 with void convection/diffusion/pressure (momentum_RK3.cpp:70-326) and
// benchmark_vortex3D::start (benchmark_vortex3D.cpp:67-101)
void vortex_step(orc_domain* d, double* phi, double* u, double* v, double* w, double* ro, double* visc,
                 double simtime, double* vol)
{
    const long n = d->sx * d->sy * d->sz;
    const double dt = d->prm.dt;
    phi_rk3(d, phi, u, v, w);
    reini(d, phi, 3);
    fluid_update(d, phi, ro, visc, vol);
    fluid_update(d, phi, ro, visc, vol);
    // momentum_RK3 with void convection, diffusion and pressure: per component
    // F = 0 + g (isource zeroes it, irhs adds gravity), then the three RK stages, the
    // zero body forcing and the ghost fill of momentum_RK3.cpp:70-326.  The benchmark
    // below overwrites u, v, w inside and refills u, v ghosts but not w's, so w's ghost
    // cells keep what this fill wrote.
    {
        double* comp[4] = {nullptr, u, v, w};
        for(int cpt = 1; cpt <= 3; ++cpt)
        {
            if(cpt == 2 && d->jdir == 0) continue;
            double* q = comp[cpt];
            int gcv = cpt == 1 ? 10 : cpt == 2 ? 11 : 12;
            double g = cpt == 1 ? d->prm.gi : cpt == 2 ? d->prm.gj : d->prm.gk;
            const double Fv = 0.0 + (0.0 + g + 0.0 + 0.0) * 1.0;
            std::vector<double> a1(q, q + n), a2(q, q + n);
            vel_loop(d, cpt, [&](long c, int, int, int) { a1[c] = q[c] + dt * 1.0 * Fv; });
            vel_loop(d, cpt, [&](long c, int, int, int) { a1[c] += 1.0 * dt * 1.0 * 0.0; });
            vel_loop(d, cpt, [&](long c, int, int, int) { a2[c] = 0.75 * q[c] + 0.25 * a1[c] + 0.25 * dt * 1.0 * Fv; });
            vel_loop(d, cpt, [&](long c, int, int, int) { a2[c] += 0.25 * dt * 1.0 * 0.0; });
            vel_loop(d, cpt, [&](long c, int, int, int) {
                q[c] = (1.0 / 3.0) * q[c] + (2.0 / 3.0) * a2[c] + (2.0 / 3.0) * dt * 1.0 * Fv;
            });
            vel_loop(d, cpt, [&](long c, int, int, int) { q[c] += (2.0 / 3.0) * dt * 1.0 * 0.0; });
            ghost(d, q, cpt, gcv);
        }
    }
    // benchmark_vortex3D::start: prescribed field at simtime (before the increment)
    const double DXM = d->DXM;
    vel_loop(d, 1, [&](long c, int i, int j, int k) {
        double xc = d->P[0][i + MARGE] + 0.5 * DXM, yc = d->P[1][j + MARGE], zc = d->P[2][k + MARGE];
        u[c] = 2.0 * pow(sin(PI_REF * xc), 2.0) * sin(2.0 * PI_REF * yc) * sin(2.0 * PI_REF * zc) * cos((PI_REF * simtime) / 3.0);
    });
    vel_loop(d, 2, [&](long c, int i, int j, int k) {
        double xc = d->P[0][i + MARGE], yc = d->P[1][j + MARGE] + 0.5 * DXM, zc = d->P[2][k + MARGE];
        v[c] = -sin(2.0 * PI_REF * xc) * pow(sin(PI_REF * yc), 2.0) * sin(2.0 * PI_REF * zc) * cos((PI_REF * simtime) / 3.0);
    });
    vel_loop(d, 3, [&](long c, int i, int j, int k) {
        double xc = d->P[0][i + MARGE], yc = d->P[1][j + MARGE], zc = d->P[2][k + MARGE] + 0.5 * DXM;
        w[c] = -sin(2.0 * PI_REF * xc) * sin(2.0 * PI_REF * yc) * pow(sin(PI_REF * zc), 2.0) * cos((PI_REF * simtime) / 3.0);
    });
    ghost(d, u, 1, 10);
    ghost(d, v, 2, 11);
    ghost(d, v, 2, 12);
}

}  // namespace

// ------------------------------------------------------------------ C API
extern "C" {

orc_domain* orc_create(int nx, int ny, int nz, int jdir, const double* XN, const double* YN, const double* ZN,
                       double DYM, double DZM, const int* group7, const orc_params* prm)
{
    orc_domain* d = new orc_domain();
    d->nx = nx; d->ny = ny; d->nz = nz; d->jdir = jdir;
    d->sx = nx + 6; d->sy = ny + 6; d->sz = nz + 6;
    for(int s = 0; s < 7; ++s) d->group[s] = group7[s];
    d->prm = *prm;
    const double* src[3] = {XN, YN, ZN};
    for(int a = 0; a < 3; ++a)
    {
        int n = d->nn(a);
        d->X[a].assign(n + 1 + 4 * MARGE, 0.0);
        for(int i = 0; i < n + 1 + 2 * MARGE; ++i) d->X[a][i] = src[a][i];
    }
    d->DYM = DYM;
    d->DZM = DZM;
    gridspacing(d);
    build_tables(d);
    // inipsi (ini_psi.cpp:28-36)
    if(jdir == 0) d->psi = prm->F45 * (1.0 / 2.0) * (d->DXM + d->DZM);
    else d->psi = prm->F45 * (1.0 / 3.0) * (d->DXM + d->DYM + d->DZM);
    d->ydir = jdir == 1 ? 1.0 : 0.0;
    d->cellnum = (long)nx * ny * nz;
    d->last[0] = 0; d->last[1] = d->last[2] = d->last[3] = 1;   // lexer::vellast, no neighbours
    build_lists(d);
    return d;
}

void orc_destroy(orc_domain* d) { delete d; }

long orc_field_size(const orc_domain* d) { return d->sx * d->sy * d->sz; }

void orc_grid_array(const orc_domain* d, int which, int axis, double* out)
{
    const std::vector<double>& v = which == 0 ? d->P[axis] : which == 1 ? d->DN[axis] : d->DP[axis];
    int n = d->nn(axis) + 2 * MARGE;
    for(int i = 0; i < n; ++i) out[i] = v[i];
}

void orc_tables(const orc_domain* d, int axis, double* qf, double* cf, double* isf)
{
    int n = d->nn(axis);
    for(int ip = 4; ip < n + 6; ++ip)
    {
        for(int r = 0; r < 24; ++r) qf[(ip - 4) * 24 + r] = d->qf[axis][(size_t)ip * 24 + r];
        for(int r = 0; r < 12; ++r) cf[(ip - 4) * 12 + r] = d->cf[axis][(size_t)ip * 12 + r];
        for(int r = 0; r < 36; ++r) isf[(ip - 4) * 36 + r] = d->isf[axis][(size_t)ip * 36 + r];
    }
}

void orc_scalars(const orc_domain* d, double* out)
{
    out[0] = d->DXM; out[1] = d->DYM; out[2] = d->DZM; out[3] = d->psi; out[4] = double(d->cellnum);
}

void orc_ghost(orc_domain* d, double* f, int kind, int gcv) { ghost(d, f, kind, gcv); }

void orc_hj(orc_domain* d, const double* phi, const double* u, const double* v, const double* w, double* L)
{
    hj(d, phi, u, v, w, L);
}

void orc_flux(orc_domain* d, int ipol, const double* q, const double* u, const double* v, const double* w, double* acc)
{
    flux(d, ipol, q, u, v, w, acc);
}

void orc_reini(orc_domain* d, double* phi, int iters) { reini(d, phi, iters); }

void orc_reini_gcv(orc_domain* d, double* phi, int iters, int gcv) { reini(d, phi, iters, gcv); }

void orc_fluid_update(orc_domain* d, const double* phi, double* ro, double* visc, double* vol2)
{
    fluid_update(d, phi, ro, visc, vol2);
}

void orc_poisson(orc_domain* d, const double* phi, const double* pcorr, double* M7, double* rhs)
{
    poisson(d, phi, pcorr, M7, rhs);
}

int orc_bicgstab(orc_domain* d, double* x, const double* M7, const double* rhs, double stop, int maxiter, double* residual)
{
    return bicgstab(d, x, M7, rhs, stop, maxiter, residual);
}

int orc_bicgstab_var(orc_domain* d, int var, double* x, const double* M7, const double* rhs, double stop, int maxiter,
                     double* residual)
{
    if(var < 1 || var > 5) return -1;
    return bicgstab(d, x, M7, rhs, stop, maxiter, residual, var);
}

long orc_veclen(orc_domain* d, int var) { return var < 1 || var > 5 ? -1 : veclen(d, var); }

void orc_idiff_assemble(orc_domain* d, int comp, double alpha, const double* vin, const double* u, const double* v,
                        const double* w, const double* au, const double* av, const double* visc, double* M7, double* rhs)
{
    idiff_assemble(d, comp, alpha, vin, u, v, w, au, av, visc, M7, rhs);
}

int orc_idiff(orc_domain* d, int comp, double alpha, double* diff, const double* vin, const double* u, const double* v,
              const double* w, const double* au, const double* av, const double* visc)
{
    return idiff(d, comp, alpha, diff, vin, u, v, w, au, av, visc);
}

int orc_set_control(orc_domain* d, const char* key, double value)
{
    if(std::string(key) == "F49" && (value == 0.0 || value == 1.0))
    {
        d->F49 = (int)value;
        return 0;
    }
    if(std::string(key) == "N43" && value > 0.0)
    {
        d->N43 = value;
        return 0;
    }
    if(std::string(key) == "df_set" && (value == 0.0 || value == 1.0))
    {
        d->df_set = (int)value;
        return 0;
    }
    if(std::string(key) == "D20" && (value == 0.0 || value == 1.0 || value == 2.0))
    {
        d->D20 = (int)value;
        return 0;
    }
    if(std::string(key) == "threads" && value >= 1.0 && value <= 256.0)
    {
        d->threads = (int)value;
        return 0;
    }
    if(std::string(key) == "dot_ext" && (value == 0.0 || value == 1.0))
    {
        d->dot_ext = (int)value;
        return 0;
    }
    return -1;
}

double orc_residual(orc_domain* d, const double* x, const double* M7, const double* rhs, double* r)
{
    return residual(d, x, M7, rhs, r);
}

int orc_pjm(orc_domain* d, const double* phi, double* u, double* v, double* w, double* press, double* pcorr, double alpha)
{
    return pjm(d, phi, u, v, w, press, pcorr, alpha);
}

int orc_fc3_step(orc_domain* d, double* phi, double* u, double* v, double* w, double* press, double* ro, double* visc,
                 double* vol2)
{
    if(d->wave.on)
    {
        // driver::loop_cfd before momentum (driver_loop_cfd.cpp:78-92): the wave precalc at
        // the step's simtime, relaxation of u, v, w, levelset_void::update (the fluid update)
        wave_precalc(d);
        wave_vel_relax(d, 1, phi, u);
        wave_vel_relax(d, 2, phi, v);
        wave_vel_relax(d, 3, phi, w);
        double vol[2];
        fluid_update(d, phi, ro, visc, vol);
    }
    FC3 s(d);
    const int it = s.step(phi, u, v, w, press, ro, visc, vol2);
    if(d->wave.on) d->simtime += d->prm.dt;
    return it;
}

void orc_wave_set(orc_domain* d, const orc_wave_desc* w, const double* phi)
{
    wave_setup(d, *w);
    // initialize::inipsi with B90 > 0 (ini_psi.cpp:40-56): psi from the mean DZN of the
    // cells within 5 DZM of the interface, in LOOP order
    int count = 0;
    double psim = 0.0;
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
        if(fabs(phi[d->id(i, j, k)]) < 5.0 * d->DZM)
        {
            psim += d->DN[2][k + MARGE];
            ++count;
        }
    d->psi = d->prm.F45 * psim / double(count);
}

void orc_wave_params(const orc_domain* d, double* out5)
{
    const orc_domain::Wave& W = d->wave;
    out5[0] = W.wT;
    out5[1] = W.wL;
    out5[2] = W.wk;
    out5[3] = W.ww;
    out5[4] = W.wa;
}

void orc_set_simtime(orc_domain* d, double t) { d->simtime = t; }

void orc_wave_relax(orc_domain* d, double simtime, const double* phi, const double* ro, double* u, double* v, double* w,
                    double* press, double* f)
{
    d->simtime = simtime;
    wave_precalc(d);
    wave_relax_uvwp(d, phi, ro, u, v, w, press);
    wave_phi_relax(d, f);
}

void orc_vortex_step(orc_domain* d, double* phi, double* u, double* v, double* w, double* ro, double* visc,
                     double simtime, double* vol2)
{
    vortex_step(d, phi, u, v, w, ro, visc, simtime, vol2);
}

void orc_phi_rk3(orc_domain* d, double* phi, const double* u, const double* v, const double* w)
{
    phi_rk3(d, phi, u, v, w);
}

// ietimestep::start (ietimestep.cpp:38-207) with N50 = 1 and ietimestep::ini
// (ietimestep.cpp:209-304): the adaptive time step of N48 = 1 (driver_logic_cfd.cpp:59-66).
// Inflow velocities W10..W16, the floating-body and fixed velocities (ufbmax, X210) are
// zero on the box domains here; maxF/G/H are zero without the EARSM closure, their only
// writer (EARSM.cpp:175-192).  out4 = dt, umax, vmax, wmax.
void orc_timestep(orc_domain* d, const double* u, const double* v, const double* w, double N47, int ini,
                  double* out4)
{
    auto MAXd = [](double a, double b) { return a > b ? a : b; };
    auto MINd = [](double a, double b) { return a < b ? a : b; };
    auto MIN3d = [&](double a, double b, double c) { return MINd(MINd(a, b), c); };
    const double W = 0.0;                 // W11_u .. W16_w, W10, ufbmax, X210_u ...
    double umax = ini ? -1.0e19 : 0.0, vmax = umax, wmax = umax;
    if(ini) umax = MAXd(W, umax);         // W10
    umax = MAXd(W, umax);                 // W11_u .. W16_u
    vmax = MAXd(W, vmax);
    wmax = MAXd(W, wmax);
    for(int i = 0; i < d->nx - d->last[1]; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
        if(d->flag[1][d->id(i, j, k)] > 0) umax = MAXd(umax, fabs(u[d->id(i, j, k)]));
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny - d->last[2]; ++j)
    for(int k = 0; k < d->nz; ++k)
        if(d->flag[2][d->id(i, j, k)] > 0) vmax = MAXd(vmax, fabs(v[d->id(i, j, k)]));
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz - d->last[3]; ++k)
        if(d->flag[3][d->id(i, j, k)] > 0) wmax = MAXd(wmax, fabs(w[d->id(i, j, k)]));
    if(ini)
    {
        umax = MAXd(umax, 2.0 * W);       // 2 ufbmax, 2 vfbmax, 2 wfbmax, 2 X210_u/v/w
        umax = MAXd(umax, 2.0);
    }
    else
    {
        umax = MAXd(umax, W);             // ufbmax, vfbmax, wfbmax
        vmax = MAXd(vmax, W);
        wmax = MAXd(wmax, W);
    }
    const double maxFGH = 0.0;
    double cu = 1.0e10;
    for(int i = 0; i < d->nx; ++i)
    for(int j = 0; j < d->ny; ++j)
    for(int k = 0; k < d->nz; ++k)
    {
        if(d->flag[4][d->id(i, j, k)] <= 0) continue;
        const double dx = MIN3d(d->DN[0][i + MARGE], d->DN[1][j + MARGE], d->DN[2][k + MARGE]);
        cu = MINd(cu, 2.0 / ((sqrt(umax * umax + vmax * vmax + wmax * wmax)) / dx + sqrt((4.0 * fabs(maxFGH)) / dx)));
    }
    cu = MIN3d(cu, 1.0e10, 1.0e10);
    double dt = N47 * cu;
    if(ini)
    {
        dt = N47 * cu * 0.25;
        dt = MAXd(dt, 1.0e-6);
    }
    out4[0] = dt;
    out4[1] = umax;
    out4[2] = vmax;
    out4[3] = wmax;
}

}  // extern "C"
