// Wave generation and absorption by relaxation zones (REEF3D's iowave with B98 = 2 and the
// numerical beach B99 = 1, 2) for linear waves (wave_lib_linear, B92 = 2): the host-side
// set-up of one tank and the device view its kernels read.
#pragma once

#include <string>
#include <vector>

#include "reef_internal.h"

namespace reef {

// per-column arrays of a tank's wave set-up, column (i, j) at i + nx * j for i in [0, nx),
// j in [0, ny); every column of the tank has an entry
enum WaveColumn
{
    WC_GEN = 0,     // 1: the cell centre lies in the generation zone (distgen < 1e20)
    WC_BEACH,       // 1: the cell centre lies in the beach zone (distbeach < 1e20)
    WC_R1WG,        // relaxation factors of the generation zone on the u, v and cell grids
    WC_R2WG,
    WC_R4WG,
    WC_R1NB,        // and of the beach
    WC_R2NB,
    WC_R4NB,
    WC_XG4,         // distances to the generation line along the wave (xgen, xgen1, xgen2):
    WC_XG1,         // the x of the wave phase at the cell centre, the u and the v point
    WC_XG2,
    WC_N
};

// device view (part of TankDev): col null = no wave generator on the tank
struct WaveDev
{
    const double* col;          // WC_N arrays of nx * ny doubles
    double* eta;                // nx * ny: the surface elevation of the last precalc
    const double* zn;           // z of the nodes (ZN, MARGE-offset): the w points' heights
    double wa, wk, ww, wdt, wD, pshift, cosg, sing;
    double phimean, psi, epsi;  // still-water level, iowave's own band half-widths
    double W50, Ui;             // air inflow velocity (W50), inflow velocity (Ui = 0 here)
    double gabs;                // |W22| of the beach's hydrostatic pressure
    double ramp_T;              // length of the ramp (B102 wT for B101 = 1, B102 for 2)
    int gen;                    // B98 == 2 (relaxation generation)
    int beach;                  // beach relaxation of u, v, w, p (B99 = 1, 2)
    int beach_phi;              // B99 == 2: the beach also relaxes phi
    int ramp_mode;              // B101
};

struct WaveSetup
{
    WaveDev dev;                // scalars (pointers filled on upload)
    double wT, wL, wH;          // wave period, length and height
    std::vector<double> col;    // WC_N * nx * ny
};

// the set-up iowave::iowave + wave_lib_parameters + wavegen_precalc_relax_func compute
// (iowave.cpp:32-263, wave_lib_parameters.cpp:28-160, iowave_dist_ini.cpp:27-139,
// iowave_precalc_relax_func.cpp:27-101); empty string or the reason it is refused
std::string build_wave(const HostGrid& h, const reef_wave_desc& w, WaveSetup& out);

}  // namespace reef
