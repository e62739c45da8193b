// REEF3D case files at the boundary: the DIVEMesh grid file and the hot-path keys of
// ctrl.txt, so a context can be built from a REEF3D case directory.  Host code only.
#include "../../include/reef3d_hip.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

thread_local std::string s_err;

int fail(int code, const std::string& m)
{
    s_err = m;
    return code;
}

constexpr int MARGE = 5;

}  // namespace

struct reef_grid_file
{
    std::vector<double> XN, YN, ZN;
};

extern "C" {

const char* reef_case_last_error(void) { return s_err.c_str(); }

// lexer::read_grid (read_grid.cpp:29-560) for one single-rank box domain: the header
// (:76-295), flag4 (:391-398), the node coordinates (:401-418), then the boundary entries
// (gcb4, 5 ints each, :441-466) from which each side's group is taken
int reef_grid_read(const char* path, reef_grid_file** out, reef_grid_desc* desc)
{
    if(!path || !out || !desc) return fail(REEF_E_ARG, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if(!f) return fail(REEF_E_ARG, std::string("cannot open grid file ") + path);
    bool ok = true;
    auto gi = [&]() {
        int v = 0;
        if(!f.read((char*)&v, sizeof v)) ok = false;
        return v;
    };
    auto gd = [&]() {
        double v = 0.0;
        if(!f.read((char*)&v, sizeof v)) ok = false;
        return v;
    };
    const int M10 = gi();
    const int nx = gi(), ny = gi(), nz = gi();
    gd();                                         // dx
    gd();                                         // DXM
    const double DYM = gd(), DZM = gd();
    for(int q = 0; q < 12; ++q) gd();             // origin, end, global min / max
    for(int q = 0; q < 6; ++q) gi();              // gknox..gknoz, origin_i..k
    const int nwall = gi();                       // gcwall_count = gcb4_count (:306)
    int npara = 0;
    for(int q = 0; q < 20; ++q) npara += gi() != 0;   // para, paraco, slpara, slparaco counts
    int nb_inner = 0;
    for(int q = 0; q < 6; ++q) nb_inner += gi() >= 0;  // nb1..nb6: neighbour ranks
    for(int q = 0; q < 6; ++q) gi();              // mx my mz + 3
    for(int q = 0; q < 6; ++q) gi();              // bcside1..6
    int periodic = 0;
    for(int q = 0; q < 9; ++q) periodic += gi() != 0;  // periodic1..3, periodicX1..6
    const int idir = gi(), jdir = gi(), kdir = gi();
    gi();                                         // P150
    const int solidread = gi(), toporead = gi();
    for(int q = 0; q < 10; ++q) gi();             // gcb estimates, cms_flag and spares
    for(int q = 0; q < 6; ++q) gd();              // global_orig_x/y, alpha_grid, spares
    if(!ok) return fail(REEF_E_ARG, "grid file header is truncated");
    if(M10 > 1 || nb_inner > 0) return fail(REEF_E_UNSUPPORTED, "grid of one rank of a parallel run");
    if(npara > 0 || periodic > 0) return fail(REEF_E_UNSUPPORTED, "grid with parallel or periodic boundaries");
    if(solidread != 0 || toporead != 0) return fail(REEF_E_UNSUPPORTED, "grid with solids or topography");
    if(idir != 1 || kdir != 1 || (jdir != 0 && jdir != 1)) return fail(REEF_E_UNSUPPORTED, "grid direction flags");
    if(nx < 1 || ny < 1 || nz < 1 || (long)nx * ny * nz > (1L << 31)) return fail(REEF_E_ARG, "bad grid size");
    for(long q = 0; q < (long)nx * ny * nz; ++q)
        if(gi() <= 0 && ok) return fail(REEF_E_UNSUPPORTED, "grid with solid cells (flag4 <= 0)");
    reef_grid_file g;
    g.XN.resize(nx + 1 + 2 * MARGE);
    g.YN.resize(ny + 1 + 2 * MARGE);
    g.ZN.resize(nz + 1 + 2 * MARGE);
    for(double& x : g.XN) x = gd();
    for(double& y : g.YN) y = gd();
    for(double& z : g.ZN) z = gd();
    int group[7] = {0, 0, 0, 0, 0, 0, 0};
    for(int q = 0; q < nwall; ++q)
    {
        gi();
        gi();
        gi();
        const int side = gi(), grp = gi();
        if(side < 1 || side > 6) continue;
        if(group[side] != 0 && group[side] != grp)
            return fail(REEF_E_UNSUPPORTED, "side " + std::to_string(side) + " mixes boundary groups");
        group[side] = grp;
    }
    if(!ok) return fail(REEF_E_ARG, "grid file is truncated");
    reef_grid_file* h = new reef_grid_file(std::move(g));
    desc->knox = nx;
    desc->knoy = ny;
    desc->knoz = nz;
    desc->j_dir = jdir;
    desc->XN = h->XN.data();
    desc->YN = h->YN.data();
    desc->ZN = h->ZN.data();
    desc->DYM = DYM;
    desc->DZM = DZM;
    for(int s = 0; s < 7; ++s) desc->group[s] = group[s];
    *out = h;
    return 0;
}

void reef_grid_file_free(reef_grid_file* file) { delete file; }

// The hot-path keys of ctrl.txt (control::read_control, read_control.cpp:28-70: a key
// letter, its number and its value; '/' starts a comment), over the defaults of ini.cpp.
// Keys that choose another scheme than the hot path's are refused.
int reef_ctrl_read(const char* path, reef_params* prm, double* ctrl4)
{
    if(!path || !prm) return fail(REEF_E_ARG, "null argument");
    std::ifstream f(path);
    if(!f) return fail(REEF_E_ARG, std::string("cannot open control file ") + path);
    // ini.cpp:389-865
    reef_params p;
    p.W1 = 998.2;
    p.W2 = 1.004e-6;
    p.W3 = 1.205;
    p.W4 = 1.41e-5;
    p.W20 = p.W21 = p.W22 = 0.0;
    p.F43 = 0.55;
    p.F45 = 2.1;
    p.dt = 1.0;          // N49
    p.N44 = 1.0e-8;
    p.N46 = 250;
    double F49 = 1, N48 = 1, N47 = 0.3, N40 = 3, D10 = 4, D20 = 2, D30 = 1, F30 = 3, F35 = 5, F40 = 3, N10 = 14,
           M10 = 0;
    std::string line;
    while(std::getline(f, line))
    {
        std::istringstream in(line);
        char c = 0;
        if(!(in >> c) || c == '/') continue;
        int num = 0;
        double v = 0.0;
        if(!(in >> num)) continue;
        if(!(in >> v)) continue;               // keys with several values: not on the path
        switch(c)
        {
            case 'W':
                if(num == 1) p.W1 = v;
                if(num == 2) p.W2 = v;
                if(num == 3) p.W3 = v;
                if(num == 4) p.W4 = v;
                if(num == 20) p.W20 = v;
                if(num == 21) p.W21 = v;
                if(num == 22) p.W22 = v;
                break;
            case 'F':
                if(num == 43) p.F43 = v;
                if(num == 45) p.F45 = v;
                if(num == 49) F49 = v;
                if(num == 30) F30 = v;
                if(num == 35) F35 = v;
                if(num == 40) F40 = v;
                break;
            case 'N':
                if(num == 44) p.N44 = v;
                if(num == 46) p.N46 = (int)v;
                if(num == 47) N47 = v;
                if(num == 48) N48 = v;
                if(num == 49) p.dt = v;
                if(num == 40) N40 = v;
                if(num == 10) N10 = v;
                break;
            case 'D':
                if(num == 10) D10 = v;
                if(num == 20) D20 = v;
                if(num == 30) D30 = v;
                break;
            case 'M':
                if(num == 10) M10 = v;
                break;
            default: break;
        }
    }
    struct Need
    {
        const char* key;
        double have;
        double want;
    };
    const Need need[] = {{"D 10 (convection: 4, weno_flux_nug)", D10, 4}, {"D 30 (pressure: 1, pjm_corr)", D30, 1},
                         {"F 30 (level set: 3)", F30, 3},           {"F 35 (level-set convection: 5)", F35, 5},
                         {"F 40 (reinit: 3, reini_RK3)", F40, 3},   {"N 10 (solver: 1, bicgstab_ijk)", N10, 1},
                         {"D 20 (diffusion: 0)", D20, 0}};
    for(const Need& n : need)
        if(n.have != n.want)
            return fail(REEF_E_UNSUPPORTED, std::string("ctrl.txt key outside the hot path: ") + n.key);
    if(N40 != 3 && N40 != 13) return fail(REEF_E_UNSUPPORTED, "ctrl.txt N 40 must be 3 (momentum_FC3) or 13");
    if(M10 > 1) return fail(REEF_E_UNSUPPORTED, "ctrl.txt M 10: one rank");
    if(F49 != 0 && F49 != 1) return fail(REEF_E_UNSUPPORTED, "ctrl.txt F 49 must be 0 or 1");
    if(N48 != 0 && N48 != 1) return fail(REEF_E_UNSUPPORTED, "ctrl.txt N 48 must be 0 or 1");
    *prm = p;
    if(ctrl4)
    {
        ctrl4[0] = F49;
        ctrl4[1] = N48;
        ctrl4[2] = N47;
        ctrl4[3] = N40;
    }
    return 0;
}

}  // extern "C"
