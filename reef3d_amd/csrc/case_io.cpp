// REEF3D case files at the boundary: the DIVEMesh grid file and the hot-path keys of
// ctrl.txt, so a context can be built from a REEF3D case directory, and the FP32 state
// file of cfd_state (Paraview and restart interop).  Host code only.
#include "../../include/reef3d_hip.h"

#include <cstdio>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <iterator>
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

// The cfd_state result file of a single-rank box (cfd_state_write.cpp:34-146): two ints
// (count, printcount) and six doubles (simtime, printtime, sedprinttime, fsfprinttime,
// probeprinttime, stateprinttime), then FP32 values in the reference's loop order, i outer
// and k inner (IJKLOOP, looping.h:129):
//   topo over ALOOP; u over ULOOP (i < knox - 1), v over VLOOP (j < knoy - 1), w over WLOOP
//   (k < knoz - 1) (looping.h:149-151 with ulast = vlast = wlast = 1, lexer::vellast for a
//   rank without neighbours or periodic sides); press, phi, kin, eps, eddyv over LOOP;
//   qbe over SLICELOOP4 (i, j); conc over LOOP.
// A box without solids has every cell in each loop.  Fields are REEF_LAYOUT_IFAST.
struct StateLayout
{
    int nx, ny, nz;
    long at(int i, int j, int k) const { return (long)(k + 3) * (ny + 6) * (nx + 6) + (long)(j + 3) * (nx + 6) + (i + 3); }
    long cells() const { return (long)nx * ny * nz; }
    // values of the staggered loops: u, v, w
    long ucells(int comp) const
    {
        return (long)(nx - (comp == 0)) * (ny - (comp == 1)) * (nz - (comp == 2));
    }
    long bytes() const
    {
        return 8 + 48 + 4 * (cells() + ucells(0) + ucells(1) + ucells(2) + 5 * cells() + (long)nx * ny + cells());
    }
    template<class F>
    void loop(int di, int dj, int dk, F f) const
    {
        for(int i = 0; i < nx - di; ++i)
            for(int j = 0; j < ny - dj; ++j)
                for(int k = 0; k < nz - dk; ++k) f(at(i, j, k));
    }
};

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
int reef_ctrl_read(const char* path, reef_params* prm, double* ctrl5)
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
                         {"F 40 (reinit: 3, reini_RK3)", F40, 3},   {"N 10 (solver: 1, bicgstab_ijk)", N10, 1}};
    for(const Need& n : need)
        if(n.have != n.want)
            return fail(REEF_E_UNSUPPORTED, std::string("ctrl.txt key outside the hot path: ") + n.key);
    if(D20 != 0 && D20 != 1 && D20 != 2)
        return fail(REEF_E_UNSUPPORTED, "ctrl.txt key outside the hot path: D 20 (diffusion: 0 diff_void, 1 ediff2, "
                                        "2 idiff2_FS)");
    if(D20 == 1 && N48 == 1)
        return fail(REEF_E_UNSUPPORTED, "ctrl.txt D 20 1 with N 48 1: the reference steps it with another "
                                        "time-step class");
    if(N40 != 3 && N40 != 13) return fail(REEF_E_UNSUPPORTED, "ctrl.txt N 40 must be 3 (momentum_FC3) or 13");
    if(M10 > 1) return fail(REEF_E_UNSUPPORTED, "ctrl.txt M 10: one rank");
    if(F49 != 0 && F49 != 1) return fail(REEF_E_UNSUPPORTED, "ctrl.txt F 49 must be 0 or 1");
    if(N48 != 0 && N48 != 1) return fail(REEF_E_UNSUPPORTED, "ctrl.txt N 48 must be 0 or 1");
    *prm = p;
    if(ctrl5)
    {
        ctrl5[0] = F49;
        ctrl5[1] = N48;
        ctrl5[2] = N47;
        ctrl5[3] = N40;
        ctrl5[4] = D20;
    }
    return 0;
}

// The values of one ctrl.txt key (read_control.cpp:28-70 reads a key letter, its number and
// the key's values), from the last line that sets it; *count = 0 when no line does.
int reef_ctrl_key(const char* path, char letter, int num, double* values, int capacity, int* count)
{
    if(!path || !count || capacity < 0 || (capacity > 0 && !values)) return fail(REEF_E_ARG, "null argument");
    std::ifstream f(path);
    if(!f) return fail(REEF_E_ARG, std::string("cannot open control file ") + path);
    *count = 0;
    std::string line;
    while(std::getline(f, line))
    {
        std::istringstream in(line);
        char c = 0;
        int n = 0;
        if(!(in >> c) || c != letter || !(in >> n) || n != num) continue;
        std::vector<double> v;
        double x;
        while(in >> x) v.push_back(x);
        const int m = (int)std::min<size_t>(v.size(), (size_t)capacity);
        for(int q = 0; q < m; ++q) values[q] = v[q];
        *count = (int)v.size();
    }
    return 0;
}

// cfd_state::write_result (cfd_state_write.cpp:34-146) of a single-rank box domain
int reef_cfd_state_write_host(int nx, int ny, int nz, const char* path, const double* const* fields,
                              const double* meta8)
{
    if(!path || !fields || !meta8) return fail(REEF_E_ARG, "null argument");
    for(int q = 0; q < 5; ++q)
        if(!fields[q]) return fail(REEF_E_ARG, "null field");
    if(nx < 1 || ny < 1 || nz < 1) return fail(REEF_E_ARG, "bad grid size");
    const StateLayout L{nx, ny, nz};
    std::vector<char> out;
    out.reserve((size_t)L.bytes());
    auto put = [&](const void* p, size_t n) { out.insert(out.end(), (const char*)p, (const char*)p + n); };
    const int head[2] = {(int)meta8[0], (int)meta8[1]};
    put(head, sizeof head);
    put(meta8 + 2, 6 * sizeof(double));
    auto putf = [&](float v) { put(&v, sizeof v); };
    L.loop(0, 0, 0, [&](long) { putf(1.0f); });                                  // topo (no topography)
    L.loop(1, 0, 0, [&](long c) { putf((float)fields[0][c]); });                 // u
    L.loop(0, 1, 0, [&](long c) { putf((float)fields[1][c]); });                 // v
    L.loop(0, 0, 1, [&](long c) { putf((float)fields[2][c]); });                 // w
    L.loop(0, 0, 0, [&](long c) { putf((float)fields[3][c]); });                 // press
    L.loop(0, 0, 0, [&](long c) { putf((float)fields[4][c]); });                 // phi
    for(int q = 0; q < 3; ++q) L.loop(0, 0, 0, [&](long) { putf(0.0f); });      // kin, eps, eddyv (laminar)
    for(long q = 0; q < (long)nx * ny; ++q) putf(0.0f);                          // qbe (no sediment)
    L.loop(0, 0, 0, [&](long) { putf(0.0f); });                                  // conc
    FILE* f = std::fopen(path, "wb");
    if(!f) return fail(REEF_E_ARG, std::string("cannot write ") + path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if(std::fclose(f) != 0 || !ok) return fail(REEF_E_ARG, std::string("write failed: ") + path);
    return 0;
}

// cfd_state::read (cfd_state_read.cpp:46-147) of a single-rank box domain: meta8 and the
// loop cells of u, v, w, press, phi (double of the FP32 value); other cells are unchanged
int reef_cfd_state_read_host(int nx, int ny, int nz, const char* path, double* const* fields, double* meta8)
{
    if(!path || !fields || !meta8) return fail(REEF_E_ARG, "null argument");
    for(int q = 0; q < 5; ++q)
        if(!fields[q]) return fail(REEF_E_ARG, "null field");
    if(nx < 1 || ny < 1 || nz < 1) return fail(REEF_E_ARG, "bad grid size");
    const StateLayout L{nx, ny, nz};
    std::ifstream f(path, std::ios::binary);
    if(!f) return fail(REEF_E_ARG, std::string("cannot open state file ") + path);
    std::vector<char> in((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if((long)in.size() != L.bytes())
        return fail(REEF_E_ARG, "state file of " + std::to_string(in.size()) + " bytes; a " + std::to_string(nx) + "x" +
                                    std::to_string(ny) + "x" + std::to_string(nz) + " box needs " +
                                    std::to_string(L.bytes()));
    const char* p = in.data();
    int head[2];
    std::memcpy(head, p, sizeof head);
    p += sizeof head;
    meta8[0] = head[0];
    meta8[1] = head[1];
    std::memcpy(meta8 + 2, p, 6 * sizeof(double));
    p += 6 * sizeof(double);
    auto getf = [&]() {
        float v;
        std::memcpy(&v, p, sizeof v);
        p += sizeof v;
        return (double)v;
    };
    p += 4 * L.cells();                                                          // topo
    L.loop(1, 0, 0, [&](long c) { fields[0][c] = getf(); });
    L.loop(0, 1, 0, [&](long c) { fields[1][c] = getf(); });
    L.loop(0, 0, 1, [&](long c) { fields[2][c] = getf(); });
    L.loop(0, 0, 0, [&](long c) { fields[3][c] = getf(); });
    L.loop(0, 0, 0, [&](long c) { fields[4][c] = getf(); });
    return 0;
}

}  // extern "C"
