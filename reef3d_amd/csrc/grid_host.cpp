// Host-side grid metadata of one box domain: spacing arrays, non-uniform WENO tables and
// the compiled ghost-cell programs.  This is what lexer/grid_helper/ghostcell hold in
// the reference (grid.cpp, weno_nug_func_precalc_*.cpp, grid_helper_*.cpp, gc_*.cpp),
// computed once per context and uploaded.
#include <algorithm>
#include "reef_internal.h"

#include <cmath>
#include <cstdio>

namespace reef {

namespace {

long round_up(long a, long b) { return (a + b - 1) / b * b; }

// Non-uniform WENO5 coefficients at index ip of coordinate array X
// (weno_nug_func_precalc_qf.cpp:27-322, _cf.cpp:27-188, _isf.cpp:27-890: the XN and XP
// sections are the same expressions on different coordinates).
void weno_coefficients(const double* X, int ip, double* q, double* c, double* s)
{
    const double a = X[ip - 2], b = X[ip - 1], x = X[ip], d = X[ip + 1], e = X[ip + 2], f = X[ip + 3],
                 h = X[ip + 4];
    // reconstruction weights of the three sub-stencils, upwind from below (0..5) / above (6..11)
    q[0] = ((e - d) / (f - x)) * ((f - d) / (e - x));
    q[1] = ((e - d) / (f - x)) * ((d - x) / (f - d));
    q[2] = ((d - x) / (e - b)) * ((d - b) / (e - x));
    q[3] = ((d - x) / (e - b)) * ((e - d) / (d - b));
    q[4] = ((d - x) / (x - a)) * ((d - b) / (d - a));
    q[5] = (1.0 + (d - x) / (d - b) + (d - x) / (d - a));
    q[6] = (1.0 + (e - d) / (f - d) + (e - d) / (h - d));
    q[7] = ((e - d) / (h - d)) * ((f - d) / (h - e));
    q[8] = ((e - d) / (f - x)) * ((f - d) / (e - x));
    q[9] = ((e - d) / (f - x)) * ((d - x) / (f - d));
    q[10] = ((d - x) / (e - b)) * ((d - b) / (e - x));
    q[11] = ((d - x) / (e - b)) * ((e - d) / (d - b));
    // optimal linear weights
    c[0] = ((d - a) / (f - a)) * ((d - b) / (f - b));
    c[1] = ((d - a) / (f - a)) * ((f - d) / (f - b)) * ((f - b) / (e - a) + 1.0);
    c[2] = ((e - d) / (f - a)) * ((f - d) / (e - a));
    c[3] = ((d - x) / (h - b)) * ((d - b) / (h - x));
    c[4] = ((d - b) / (h - b)) * ((h - d) / (h - x)) * ((h - x) / (f - b) + 1.0);
    c[5] = ((f - d) / (h - b)) * ((h - d) / (f - b));
    // smoothness-indicator quadratic forms
    double k;
    k = 4.0 * pow((d - x) / (f - x), 2.0);
    s[0] = k * ((10.0 * pow(d - x, 2.0) + (e - x) * (e - d)) / pow(f - d, 2.0));
    s[1] = k * ((20.0 * pow(d - x, 2.0) + 2.0 * (e - x) * (e - d) + (f - x) * (2.0 * e - d - x)) / ((f - d) * (e - x)));
    s[2] = k * ((10.0 * pow(d - x, 2.0) + (f + e - 2.0 * x) * (f + e - d - x)) / pow(e - x, 2.0));
    k = 4.0 * pow((d - x) / (e - b), 2.0);
    s[3] = k * ((10.0 * pow(d - x, 2.0) + (e - x) * (e - d)) / pow(d - b, 2.0));
    s[4] = k * ((20.0 * pow(d - x, 2.0) - (e - d) * (x - b) - (e - x) * (d - b)) / ((e - x) * (d - b)));
    s[5] = k * ((10.0 * pow(d - x, 2.0) + (x - b) * (d - b)) / pow(e - x, 2.0));
    k = 4.0 * pow((d - x) / (d - a), 2.0);
    s[6] = k * ((10.0 * pow(d - x, 2.0) + (d - b) * (x - b)) / pow(x - a, 2.0));
    s[7] = k * ((20.0 * pow(d - x, 2.0) + 2.0 * (d - b) * (x - b) + (d - a) * (d + x - 2.0 * b)) / ((d - b) * (x - a)));
    s[8] = k * ((10.0 * pow(d - x, 2.0) + (2.0 * d - a - b) * (d + x - b - a)) / pow(d - b, 2.0));
    k = 4.0 * pow((e - d) / (h - d), 2.0);
    s[9] = k * ((10.0 * pow(e - d, 2.0) + (f - d) * (f - e)) / pow(h - e, 2.0));
    s[10] = k * ((20.0 * pow(e - d, 2.0) + 2.0 * (f - d) * (f - e) + (h - d) * (2.0 * f - e - d)) / ((h - e) * (f - d)));
    s[11] = k * ((10.0 * pow(e - d, 2.0) + (h + f - 2.0 * d) * (h + f - e - d)) / pow(f - d, 2.0));
    k = 4.0 * pow((e - d) / (f - x), 2.0);
    s[12] = k * ((10.0 * pow(e - d, 2.0) + (f - d) * (f - e)) / pow(e - x, 2.0));
    s[13] = k * ((20.0 * pow(e - d, 2.0) - (f - e) * (d - x) - (f - d) * (e - x)) / ((f - d) * (e - x)));
    s[14] = k * ((10.0 * pow(e - d, 2.0) + (d - x) * (e - x)) / pow(f - d, 2.0));
    k = 4.0 * pow((e - d) / (e - b), 2.0);
    s[15] = k * ((10.0 * pow(e - d, 2.0) + (e - x) * (d - x)) / pow(d - b, 2.0));
    s[16] = k * ((20.0 * pow(e - d, 2.0) + 2.0 * (e - x) * (d - x) + (e - b) * (e + d - 2.0 * x)) / ((e - x) * (d - b)));
    s[17] = k * ((10.0 * pow(e - d, 2.0) + (2.0 * e - b - x) * (e + d - x - b)) / pow(e - x, 2.0));
}

struct Face
{
    int i, j, k, side, group;
    double dist;
    bool active;
};

// side -> outward offset (1 -x, 2 +y, 3 -y, 4 +x, 5 -z, 6 +z)
const int SIDE_D[7][3] = {{0, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {0, 0, -1}, {0, 0, 1}};
int side_axis(int s) { return (s == 1 || s == 4) ? 0 : (s == 2 || s == 3) ? 1 : 2; }

// boundary entries of the box (read_grid.cpp:458-488 as written by the grid generator),
// LOOP order, sides 1..6; velocity component c shifts its "last" face one cell inwards
// and disables the other entries of that cell row (grid_helper_gcb1/2/3.cpp)
std::vector<Face> box_faces(const HostGrid& g, int comp)
{
    std::vector<Face> out;
    out.reserve(2 * ((size_t)g.ny * g.nz + (size_t)g.nx * g.nz + (size_t)g.nx * g.ny));
    for(int i = 0; i < g.nx; ++i)
    for(int j = 0; j < g.ny; ++j)
    for(int k = 0; k < g.nz; ++k)
    {
        const bool at[7] = {false, i == 0, j == g.ny - 1, j == 0, i == g.nx - 1, k == 0, k == g.nz - 1};
        for(int s = 1; s <= 6; ++s)
        {
            if(!at[s]) continue;
            int ax = side_axis(s);
            int idx = ax == 0 ? i : ax == 1 ? j : k;
            Face f{i, j, k, s, g.group[s], 0.5 * g.DP[ax][idx + MARGE], true};
            out.push_back(f);
        }
    }
    if(comp >= 1 && comp <= 3)
    {
        const int lastside = comp == 1 ? 4 : comp == 2 ? 2 : 6;
        const int ax = comp - 1;
        const int nlast = ax == 0 ? g.nx - 1 : ax == 1 ? g.ny - 1 : g.nz - 1;
        for(Face& f : out)
        {
            if(side_axis(f.side) == ax)
            {
                int idx = ax == 0 ? f.i : ax == 1 ? f.j : f.k;
                f.dist += 0.5 * g.DP[ax][idx + MARGE];
            }
        }
        for(Face& f : out)
        {
            int idx = ax == 0 ? f.i : ax == 1 ? f.j : f.k;
            if(f.side == lastside)
            {
                if(ax == 0) f.i -= 1;
                if(ax == 1) f.j -= 1;
                if(ax == 2) f.k -= 1;
            }
            else if(idx == nlast)
                f.active = false;   // row of the shifted face: its other entries are disabled
        }
    }
    return out;
}

// gceval1..4 restricted to groups 21/3 and the hot-path tags (gc_epol1.cpp:26-121,
// gc_epol2.cpp, gc_epol3.cpp, gc_epol4.cpp:26-317, labels of ghostcell.cpp:45-200):
// 0 untouched, 1 wall-normal extrapolation, 4 copy, 5 zero
int bc_rule(int kind, int gcv, int group, int side)
{
    if(gcv == 50) return 4;
    if(kind == 4)
    {
        if(gcv == 1) return 4;
        if((gcv >= 51 && gcv <= 54) || gcv == 40) return (group == 21 || group == 3) ? 4 : 0;
        return 0;
    }
    const bool normal = side_axis(side) == kind - 1;
    const int tag_main = 9 + kind, tag_corr = 6 + kind, tag_neu = 13 + kind, tag_fs = 16 + kind;
    const bool vel = gcv == tag_main || gcv == kind;
    if(group == 21)
    {
        if(vel) return normal ? 1 : 5;
        if(gcv == tag_neu) return 4;
        if(gcv == tag_corr && normal) return 5;
        return 0;
    }
    if(group == 3 && (vel || gcv == tag_fs))
        return normal ? (kind == 3 ? 5 : 1) : 4;
    return 0;
}

}  // namespace

std::string build_grid(HostGrid& g, const reef_grid_desc& d, double F45)
{
    if(d.knox < 4 || d.knoy < 1 || d.knoz < 4) return "grid too small (need knox, knoz >= 4)";
    if(d.j_dir == 0 && d.knoy != 1) return "2-D domain needs knoy == 1";
    if(d.j_dir == 1 && d.knoy < 4) return "3-D domain needs knoy >= 4";
    for(int s = 1; s <= 6; ++s)
    {
        if(d.group[s] != 21 && !(s == 6 && d.group[s] == 3))
            return "only wall (21) sides and a free-surface (3) top are supported";
    }
    g.nx = d.knox; g.ny = d.knoy; g.nz = d.knoz; g.jdir = d.j_dir;
    for(int s = 0; s < 7; ++s) g.group[s] = d.group[s];
    const double* src[3] = {d.XN, d.YN, d.ZN};
    const int n3[3] = {g.nx, g.ny, g.nz};
    for(int a = 0; a < 3; ++a)
    {
        const int n = n3[a];
        const int len = n + 1 + 4 * MARGE;
        g.X[a].assign(len, 0.0);
        for(int i = 0; i < n + 1 + 2 * MARGE; ++i) g.X[a][i] = src[a][i];
        g.P[a].assign(len, 0.0);
        g.DN[a].assign(len, 0.0);
        g.DP[a].assign(len, 0.0);
        // grid::gridspacing (grid.cpp:96-124)
        for(int ip = 0; ip < n + 2 * MARGE; ++ip)
        {
            const double* X = g.X[a].data();
            g.P[a][ip] = 0.5 * (X[ip] + X[ip + 1]);
            g.DN[a][ip] = X[ip + 1] - X[ip];
            g.DP[a][ip] = 0.5 * (X[ip + 2] + X[ip + 1]) - 0.5 * (X[ip + 1] + X[ip]);
        }
        for(int ip = 1; ip < n + 2 * MARGE; ++ip)
            if(!(g.DN[a][ip] > 0.0)) return "node coordinates must increase";
    }
    // mean spacing (grid.cpp:126-149), psi (ini_psi.cpp:32-36)
    double sum = 0.0;
    int count = 0;
    for(int i = 0; i < g.nx; ++i) { sum += g.DP[0][i + MARGE]; ++count; }
    if(g.jdir == 1)
        for(int j = 0; j < g.ny; ++j) { sum += g.DP[1][j + MARGE]; ++count; }
    for(int k = 0; k < g.nz; ++k) { sum += g.DP[2][k + MARGE]; ++count; }
    g.DXM = sum / double(count);
    g.DYM = d.DYM;
    g.DZM = d.DZM;
    g.psi = g.jdir == 0 ? F45 * (1.0 / 2.0) * (g.DXM + g.DZM) : F45 * (1.0 / 3.0) * (g.DXM + g.DYM + g.DZM);
    // WENO tables for IP in [4, n+6) (IBLOOP), uf = 0 on XN, uf = 1 on XP
    for(int a = 0; a < 3; ++a)
    {
        const int n = n3[a];
        g.qf[a].assign((size_t)(n + 8) * 24, 0.0);
        g.cf[a].assign((size_t)(n + 8) * 12, 0.0);
        g.isf[a].assign((size_t)(n + 8) * 36, 0.0);
        for(int ip = MARGE - 1; ip < n + MARGE + 1; ++ip)
            for(int uf = 0; uf < 2; ++uf)
                weno_coefficients(uf == 0 ? g.X[a].data() : g.P[a].data(), ip, &g.qf[a][(size_t)ip * 24 + uf * 12],
                                  &g.cf[a][(size_t)ip * 12 + uf * 6], &g.isf[a][(size_t)ip * 36 + uf * 18]);
    }
    // device layout: rows padded to 32 elements, (0,j,k) at a multiple of 32 elements, so
    // every row starts on a 128-byte boundary in the FP32 arrays as in the FP64 ones (a
    // wavefront's 64-cell row is then 2 whole cache lines of FP32, 4 of FP64)
    g.sx = round_up(g.nx + 2 * NG, 32);
    g.sy = g.ny + 2 * NG;
    g.sz = g.nz + 2 * NG;
    g.sxy = g.sx * g.sy;
    g.base = 32 + NG * g.sx + NG * g.sxy;        // (-3, j, k) at offset 29 within a row
    g.nalloc = round_up(32 + g.sxy * g.sz + 64, 64);
    // the solver walks address an array through 32-bit byte offsets from its base (FP64:
    // at most 2^29 - 1 elements) and the ghost programs through 32-bit element indices
    if(g.nalloc * (long)sizeof(double) >= (1L << 32))
        return "field too large: at most 2^29 - 1 doubles per field (32-bit byte offsets of the solver kernels)";
    g.last[0] = 0;
    g.last[1] = g.last[2] = g.last[3] = 1;        // lexer::vellast, no neighbour ranks
    // flags (gc_flagfield.cpp:26-98): 10 inside, -20 outside; u/v/w lose their last row
    const long ncomp = (long)(g.nx + 6) * (g.ny + 6) * (g.nz + 6);
    for(int f = 1; f <= 4; ++f)
    {
        g.flag[f].assign(ncomp, -20);
        for(int k = 0; k < g.nz; ++k)
        for(int j = 0; j < g.ny; ++j)
        for(int i = 0; i < g.nx; ++i)
        {
            bool last = (f == 1 && i == g.nx - 1) || (f == 2 && j == g.ny - 1) || (f == 3 && k == g.nz - 1);
            g.flag[f][g.cid(i, j, k)] = last ? -20 : 10;
        }
    }
    return "";
}

namespace {

// one program of ghostcell::start<kind>(f, gcv); lists[1..4] are the boundary entries of
// the u, v, w and scalar grids, hgc the mark array accumulated over lists[1..kind]
std::string ghost_program(const HostGrid& g, const std::vector<Face>* lists, const std::vector<int>& hgc, int kind,
                          int gcv, std::vector<GhostOp>& ops)
{
    ops.clear();
    const long ncomp = (long)(g.nx + 6) * (g.ny + 6) * (g.nz + 6);
    std::vector<char> written(ncomp, 0), read(ncomp, 0);
    auto emit = [&](int i, int j, int k, int si, int sj, int sk, int mode, double w) -> bool {
        long dc = g.cid(i, j, k);
        if(written[dc]) return false;
        written[dc] = 1;
        GhostOp op;
        op.dst = (int32_t)g.did(i, j, k);
        op.w = w;
        if(mode == 5)
            op.src = -1;
        else
        {
            read[g.cid(si, sj, sk)] = 1;
            op.src = (int32_t)g.did(si, sj, sk) | (mode == 1 ? (1 << 30) : 0);
        }
        ops.push_back(op);
        return true;
    };
    const std::vector<Face>& faces = lists[kind];
    const std::vector<Face>& faces1 = lists[1];
    // face ghosts (gcdistro1..4, gc_start.cpp:38-40 etc.)
    if(kind != 2 || g.jdir == 1)
        for(size_t q = 0; q < faces.size(); ++q)
        {
            const Face& f = faces[q];
            if(!f.active) continue;
            // gc_start.cpp:39 tests the u list's side for every kind
            const int s1 = faces1[q].active ? faces1[q].side : -faces1[q].side;
            if(!((s1 != 2 && s1 != 3) || g.jdir == 1)) continue;
            int rule = bc_rule(kind, gcv, f.group, f.side);
            if(rule == 0) continue;
            const int ax = side_axis(f.side);
            if(ax == 1 && g.jdir == 0) continue;      // neumann / noslip skip y sides in 2-D
            const int* o = SIDE_D[f.side];
            if(rule == 4 || rule == 5)
            {
                for(int n = 1; n <= 3; ++n)
                    if(!emit(f.i + o[0] * n, f.j + o[1] * n, f.k + o[2] * n, f.i, f.j, f.k, rule, 1.0))
                        return "overlapping ghost writes";
                continue;
            }
            // gc_dirichlet-ortho.cpp:28-137, orderdir 2, wall value 0
            int idx = ax == 0 ? f.i : ax == 1 ? f.j : f.k;
            const double dx = g.DP[ax][idx + MARGE];
            int ys = 1;
            if(f.dist > g.DXM * (1.0 - 1.0e-6) && f.dist < g.DXM * (1.0 + 1.0e-6)) ys = 0;
            if(ys == 1 && f.dist < 0.5 * dx) return "image-point extrapolation not supported";
            double wq[3];
            for(int q2 = 0; q2 < 3; ++q2)
            {
                const double xq = dx * double(q2 + 2 - ys);
                wq[q2] = 1.0 * ((xq - f.dist) / (0.0 - f.dist + 1.0e-20));
            }
            for(int n = 0; n < 3; ++n)
            {
                // layer n gets y[1 + n + ys]: y[1] is the wall value, y[2+q] the extrapolants
                const int yi = 1 + n + ys;
                const int i2 = f.i + o[0] * (n + 1), j2 = f.j + o[1] * (n + 1), k2 = f.k + o[2] * (n + 1);
                bool ok = yi == 1 ? emit(i2, j2, k2, 0, 0, 0, 5, 0.0)
                                  : emit(i2, j2, k2, f.i, f.j, f.k, 1, wq[yi - 2]);
                if(!ok) return "overlapping ghost writes";
            }
        }
    // edge / corner ghosts (grid_helper_dgc1..4.cpp, gc_dgcpol1..4.cpp)
    static const int DIRS[20][3] = {
        {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
        {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}, {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
        {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
    std::vector<char> dwritten(ncomp, 0);
    {
        const int c = kind;
        const std::vector<int>& fl = g.flag[c];
        const int il = c == 1, jl = c == 2, kl = c == 3;
        const int ni = g.nx - il, nj = g.ny - jl, nk = g.nz - kl;
        // the flags are positive exactly on the loop range.  Every direction has two or
        // three nonzero components and each must lead out of the range, so only cells with
        // at least two coordinates on its rim can emit: columns with i and j on the rim
        // visit every k, columns with one of them their first and last k, others none
        for(int i = 0; i < ni; ++i)
        for(int j = 0; j < nj; ++j)
        {
            const int rim = (i == 0 || i == ni - 1) + (j == 0 || j == nj - 1);
            if(rim == 0) continue;
            const int kstep = rim == 2 ? 1 : std::max(1, nk - 1);
        for(int k = 0; k < nk; k += kstep)
        {
            if(fl[g.cid(i, j, k)] <= 0) continue;
            for(const auto& dd : DIRS)
            {
                const int i2 = i + dd[0], j2 = j + dd[1], k2 = k + dd[2];
                if(fl[g.cid(i2, j2, k2)] >= 0) continue;
                if(dd[0] && fl[g.cid(i2, j, k)] >= 0) continue;
                if(dd[1] && fl[g.cid(i, j2, k)] >= 0) continue;
                if(dd[2] && fl[g.cid(i, j, k2)] >= 0) continue;
                if(hgc[g.cid(i2, j2, k2)] != 0) continue;
                long dc = g.cid(i2, j2, k2);
                if(written[dc] || dwritten[dc]) return "overlapping edge ghost writes";
                dwritten[dc] = 1;
                read[g.cid(i, j, k)] = 1;
                GhostOp op;
                op.dst = (int32_t)g.did(i2, j2, k2);
                op.src = (int32_t)g.did(i, j, k);
                op.w = 1.0;
                ops.push_back(op);
            }
        }
        }
    }
    // every source must be a cell this program does not write: one parallel pass is exact
    for(long c = 0; c < ncomp; ++c)
        if(read[c] && (written[c] || dwritten[c])) return "ghost program reads a cell it writes";
    return "";
}

}  // namespace

std::string build_ghost_programs(const HostGrid& g, int n, const int (*kind_gcv)[2], std::vector<GhostOp>* ops)
{
    for(int q = 0; q < n; ++q)
        if(kind_gcv[q][0] < 1 || kind_gcv[q][0] > 4) return "bad ghost program kind";
    std::vector<Face> lists[5];
    for(int c = 1; c <= 4; ++c) lists[c] = box_faces(g, c == 4 ? 0 : c);
    // the four velocity/scalar lists share one mark array (grid_helper.cpp:42-59: it is
    // cleared once before fill_dgc1 and accumulates through fill_dgc4); programs are
    // compiled in increasing kind so the array only ever gains lists
    std::vector<int> order(n);
    for(int q = 0; q < n; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return kind_gcv[a][0] < kind_gcv[b][0]; });
    const long ncomp = (long)(g.nx + 6) * (g.ny + 6) * (g.nz + 6);
    std::vector<int> hgc(ncomp, 0);
    int marked = 0;
    for(int q : order)
    {
        const int kind = kind_gcv[q][0];
        for(; marked < kind; )
        {
            ++marked;
            for(const Face& f : lists[marked])
            {
                if(!f.active) continue;
                const int* o = SIDE_D[f.side];
                for(int m = 1; m <= 3; ++m) hgc[g.cid(f.i + o[0] * m, f.j + o[1] * m, f.k + o[2] * m)] += 1;
            }
        }
        std::string e = ghost_program(g, lists, hgc, kind, kind_gcv[q][1], ops[q]);
        if(!e.empty()) return e;
    }
    return "";
}

std::string build_ghost_program(const HostGrid& g, int kind, int gcv, std::vector<GhostOp>& ops)
{
    const int kg[1][2] = {{kind, gcv}};
    return build_ghost_programs(g, 1, kg, &ops);
}

}  // namespace reef
