// FP64 stencil kernels of the level-set and momentum convection: HJ-WENO5 transport
// stages of phi, the reinitialisation sub-stages, and the conservative WENO5 flux of the
// staggered velocities fused with the momentum_FC3 stage update.
//
// Every kernel reproduces the reference's per-cell arithmetic in the reference's
// operation order (the library is built with -ffp-contract=off), so results are
// bit-identical to the CPU path on the same inputs.
#include "reef_internal.h"

namespace reef {

// Linear block id -> tile, remapped so that each of the 8 XCDs works on one contiguous
// slab of the domain (workgroups are dispatched to XCDs round robin).  Grids are padded
// to a multiple of 8 blocks; padding blocks return false.
__device__ __forceinline__ bool tile_of(int nbx, int nby, int nbtot, int& bx, int& by, int& bz)
{
    const unsigned nb8 = gridDim.x;
    const unsigned per = nb8 >> 3;
    const unsigned b = blockIdx.x;
    const unsigned lin = (b & 7u) * per + (b >> 3);
    if(lin >= (unsigned)nbtot) return false;
    bz = lin / (nbx * nby);
    const unsigned rem = lin - (unsigned)bz * nbx * nby;
    by = rem / nbx;
    bx = rem - (unsigned)by * nbx;
    return true;
}

constexpr int TX = 64, TY = 4;

static inline dim3 tile_grid(const DevGrid& g, int nx, int ny, int nz, int& nbx, int& nby, int& nbtot)
{
    nbx = (nx + TX - 1) / TX;
    nby = (ny + TY - 1) / TY;
    nbtot = nbx * nby * nz;
    return dim3((unsigned)((nbtot + 7) / 8 * 8));
}

// ------------------------------------------------------------------ WENO5 core
// Smoothness indicators, (IS+psi)^-2 weights without alpha normalisation and the
// non-uniform reconstruction, in the operation order of weno_nug_func.h:48-196 and
// weno_flux_nug.cpp:136-160 / weno_hj_df_nug.cpp:110-177.
__device__ __forceinline__ double weno_recon(const double* __restrict__ I, const double* __restrict__ C,
                                             const double* __restrict__ Q, bool minus, double q1, double q2,
                                             double q3, double q4, double q5)
{
    const double dq12 = q1 - q2, dq23 = q2 - q3, dq32 = q3 - q2, dq34 = q3 - q4, dq43 = q4 - q3, dq54 = q5 - q4;
    const int r = minus ? 0 : 3;
    const double* I1 = I + r * 3;
    const double is1 = I1[0] * dq54 * dq54 + I1[1] * (dq54) * (dq34) + I1[2] * dq34 * dq34;
    const double is2 = I1[3] * dq23 * dq23 + I1[4] * (dq43) * (dq23) + I1[5] * dq43 * dq43;
    const double is3 = I1[6] * dq12 * dq12 + I1[7] * (dq32) * (dq12) + I1[8] * dq32 * dq32;
    const double a1 = is1 + 1.0e-6, a2 = is2 + 1.0e-6, a3 = is3 + 1.0e-6;
    const double c0 = C[r], c1 = C[r + 1], c2 = C[r + 2];
    const double w1 = c0 / (0.0 + (a1 * a1) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
    const double w2 = c1 / (0.0 + (a2 * a2) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
    const double w3 = c2 / (0.0 + (a3 * a3) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
    if(minus)
        return w1 * (q4 + Q[0] * (q3 - q4) - Q[1] * (q5 - q4)) + w2 * (q3 + Q[2] * (q4 - q3) - Q[3] * (q2 - q3))
             + w3 * (q2 + Q[4] * (q1 - q2) + Q[5] * (q3 - q2));
    return w1 * (q4 + Q[6] * (q3 - q4) + Q[7] * (q5 - q4)) + w2 * (q3 + Q[8] * (q2 - q3) - Q[9] * (q4 - q3))
         + w3 * (q2 + Q[10] * (q3 - q2) - Q[11] * (q1 - q2));
}

// HJ gradient along one axis at index ix (stride s), divided differences over DXP
// (weno_hj_df_nug.cpp:344-396 / ddweno_nug_sf.cpp:186-224)
__device__ __forceinline__ double hj_grad(const DevGrid& g, int axis, const double* __restrict__ f, long c, long s,
                                          int ix, double advec)
{
    const bool minus = advec > 0.0;
    if(!minus && !(advec < 0.0)) return 0.0;
    const int ip = ix + MARGE;
    const double* D = axis == 0 ? g.DXP : axis == 1 ? g.DYP : g.DZP;
    const double* I = g.isf[axis] + (size_t)ip * 36;
    const double* C = g.cf[axis] + (size_t)ip * 12;
    const double* Q = g.qf[axis] + (size_t)ip * 24;
    const double fm3 = f[c - 3 * s], fm2 = f[c - 2 * s], fm1 = f[c - s], f0 = f[c], fp1 = f[c + s], fp2 = f[c + 2 * s],
                 fp3 = f[c + 3 * s];
    if(minus)
        return weno_recon(I, C, Q, true, (fm2 - fm3) / D[ip - 3], (fm1 - fm2) / D[ip - 2], (f0 - fm1) / D[ip - 1],
                          (fp1 - f0) / D[ip], (fp2 - fp1) / D[ip + 1]);
    return weno_recon(I, C, Q, false, (fm1 - fm2) / D[ip - 2], (f0 - fm1) / D[ip - 1], (fp1 - f0) / D[ip],
                      (fp2 - fp1) / D[ip + 1], (fp3 - fp2) / D[ip + 2]);
}

// ------------------------------------------------------------------ HJ-WENO stage
// weno_hj_df_nug::aij (weno_hj_df_nug.cpp:85-108) with flux_HJ_CDS2 ipol 4 velocities,
// fused with the TVD-RK3 combination of levelset_RK3.cpp:118-157 / momentum_FC3.cpp.
__global__ void __launch_bounds__(256) k_hj_stage(DevGrid g, int nbx, int nby, int nbtot, int stage,
                                                  const double* __restrict__ phi_s, const double* __restrict__ phi_n,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, double* __restrict__ out, double cdt)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(i >= g.nx || j >= g.ny) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
    const double iadvec = 0.5 * (u[c] + u[c - 1]);
    const double jadvec = 0.5 * (v[c] + v[c - g.sx]);
    const double kadvec = 0.5 * (w[c] + w[c - g.sxy]);
    double L = -iadvec * hj_grad(g, 0, phi_s, c, 1, i, iadvec);
    if(g.jdir == 1) L -= jadvec * hj_grad(g, 1, phi_s, c, g.sx, j, jadvec);
    L -= kadvec * hj_grad(g, 2, phi_s, c, g.sxy, k, kadvec);
    const double La = 0.0 + L;     // a->L zeroed, then += aij
    if(stage == 0)
        out[c] += L;
    else if(stage == 1)
        out[c] = phi_s[c] + cdt * La;
    else if(stage == 2)
        out[c] = 0.75 * phi_n[c] + 0.25 * phi_s[c] + cdt * La;
    else
        out[c] = (1.0 / 3.0) * phi_n[c] + (2.0 / 3.0) * phi_s[c] + cdt * La;
}

void launch_hj_stage(const DevGrid& g, int stage, const double* phi_s, const double* phi_n, const double* u,
                     const double* v, const double* w, double* out, double cdt, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_hj_stage, grid, dim3(256), 0, s, g, nbx, nby, nbtot, stage, phi_s, phi_n, u, v, w, out, cdt);
}

// ------------------------------------------------------------------ reinitialisation
// reinidisc_f::disc (reinidisc_f.cpp:57-127) fused with the reini_RK3 sub-stage
// (reini_RK3.cpp:114-139) and the per-cell pseudo time step (reini_RK3.cpp:150-160).
// fsf = 1 is reinidisc_fsf::disc (reinidisc_fsf.cpp:66-141, F49 = 0): the same update
// on cells whose 7-point stencil has one sign, L = 0 elsewhere (ddweno_nug and
// ddweno_nug_sf in modus 0 compute the same gradients).
__global__ void __launch_bounds__(256) k_reini_stage(DevGrid g, int nbx, int nby, int nbtot, int stage,
                                                     const double* __restrict__ f, const double* __restrict__ f0,
                                                     double* __restrict__ out, double F43, int fsf)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(i >= g.nx || j >= g.ny) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
    const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
    const double lsv = f[c];
    bool active = true;
    if(fsf)
    {
        const double n1 = f[c + 1], n2 = f[c - 1], n3 = f[c + g.sx], n4 = f[c - g.sx], n5 = f[c + g.sxy],
                     n6 = f[c - g.sxy];
        active = (lsv >= 0.0 && n1 >= 0.0 && n2 >= 0.0 && n3 >= 0.0 && n4 >= 0.0 && n5 >= 0.0 && n6 >= 0.0)
              || (lsv < 0.0 && n1 < 0.0 && n2 < 0.0 && n3 < 0.0 && n4 < 0.0 && n5 < 0.0 && n6 < 0.0);
    }
    double L = 0.0;
    if(active)
    {
    double dx = 0.0, dy = 0.0, dz = 0.0;
    double lsSig = lsv / sqrt(lsv * lsv);
    if(fabs(lsv) < 1.0e-8) lsSig = 1.0;
    const double xmin = (lsv - f[c - 1]) / g.DXP[ip - 1];
    const double xplus = (f[c + 1] - lsv) / g.DXP[ip];
    if(xmin * lsSig > 0.0 && xplus * lsSig > -xmin * lsSig) dx = hj_grad(g, 0, f, c, 1, i, 1.0);
    if(xplus * lsSig < 0.0 && xmin * lsSig < -xplus * lsSig) dx = hj_grad(g, 0, f, c, 1, i, -1.0);
    if(xplus * lsSig > 0.0 && xmin * lsSig < 0.0) dx = 0.0;
    if(g.jdir == 1)
    {
        const double ymin = (lsv - f[c - g.sx]) / g.DYP[jp - 1];
        const double yplus = (f[c + g.sx] - lsv) / g.DYP[jp];
        if(ymin * lsSig > 0.0 && yplus * lsSig > -ymin * lsSig) dy = hj_grad(g, 1, f, c, g.sx, j, 1.0);
        if(yplus * lsSig < 0.0 && ymin * lsSig < -yplus * lsSig) dy = hj_grad(g, 1, f, c, g.sx, j, -1.0);
        if(yplus * lsSig > 0.0 && ymin * lsSig < 0.0) dy = 0.0;
    }
    const double zmin = (lsv - f[c - g.sxy]) / g.DZP[kp - 1];
    const double zplus = (f[c + g.sxy] - lsv) / g.DZP[kp];
    if(zmin * lsSig > 0.0 && zplus * lsSig > -zmin * lsSig) dz = hj_grad(g, 2, f, c, g.sxy, k, 1.0);
    if(zplus * lsSig < 0.0 && zmin * lsSig < -zplus * lsSig) dz = hj_grad(g, 2, f, c, g.sxy, k, -1.0);
    if(zplus * lsSig > 0.0 && zmin * lsSig < 0.0) dz = 0.0;
    const double dnorm = sqrt(dx * dx + dy * dy + dz * dz);
    double deltax;
    if(g.jdir == 0)
        deltax = (1.0 / 2.0) * (g.DXN[ip] + g.DZN[kp]);
    else
        deltax = (1.0 / 3.0) * (g.DXN[ip] + g.DYN[jp] + g.DZN[kp]);
    double sign = lsv / sqrt(lsv * lsv + dnorm * dnorm * deltax * deltax);
    if(sign != sign) sign = 1.0;
    L = -(sign * dnorm - sign);
    }
    // pseudo time step F43 * min(DXP, DYP, DZP)
    const double a = g.DXP[ip], b = g.DYP[jp], cc = g.DZP[kp];
    const double m = g.jdir == 0 ? (a < cc ? a : cc) : (((a < b ? a : b) < cc) ? (a < b ? a : b) : cc);
    const double dtc = F43 * m;
    if(stage == 1)
        out[c] = f0[c] + dtc * L;
    else if(stage == 2)
        out[c] = 0.75 * f0[c] + 0.25 * lsv + 0.25 * dtc * L;
    else
        out[c] = (1.0 / 3.0) * f0[c] + (2.0 / 3.0) * lsv + (2.0 / 3.0) * dtc * L;
}

void launch_reini_stage(const DevGrid& g, int stage, const double* fs, const double* f0, double* out, double F43,
                        int fsf, hipStream_t s)
{
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_reini_stage, grid, dim3(256), 0, s, g, nbx, nby, nbtot, stage, fs, f0, out, F43, fsf);
}

// ------------------------------------------------------------------ WENO flux
// weno_flux_nug::fx/fy/fz with iqmin..kqmax (weno_flux_nug.cpp:136-313): values; the
// x upwind-from-above stencil repeats f(i+2) (weno_flux_nug.cpp:288-295).
__device__ __forceinline__ double flux_face(const DevGrid& g, int axis, int uf, const double* __restrict__ f, long c,
                                            long s, int ix, double advec)
{
    const bool minus = advec > 0.0;
    if(!minus && !(advec < 0.0)) return 0.0;
    const int ip = ix + MARGE;
    const double* I = g.isf[axis] + (size_t)ip * 36 + uf * 18;
    const double* C = g.cf[axis] + (size_t)ip * 12 + uf * 6;
    const double* Q = g.qf[axis] + (size_t)ip * 24 + uf * 12;
    if(minus) return weno_recon(I, C, Q, true, f[c - 2 * s], f[c - s], f[c], f[c + s], f[c + 2 * s]);
    return weno_recon(I, C, Q, false, f[c - s], f[c], f[c + s], f[c + 2 * s], axis == 0 ? f[c + 2 * s] : f[c + 3 * s]);
}

// flux divergence of component ipol at cell c (weno_flux_nug.cpp:110-134) with the
// flux_face_CDS2 face velocities (flux_face_CDS2.cpp:36-116)
__device__ __forceinline__ double flux_div(const DevGrid& g, int ipol, const double* __restrict__ q,
                                           const double* __restrict__ u, const double* __restrict__ v,
                                           const double* __restrict__ w, long c, int i, int j, int k)
{
    const long sx = g.sx, sxy = g.sxy;
    double iv1, iv2, jv1, jv2, kv1, kv2;
    if(ipol == 1)
    {
        iv1 = 0.5 * (u[c] + u[c - 1]);            iv2 = 0.5 * (u[c] + u[c + 1]);
        jv1 = 0.5 * (v[c - sx] + v[c + 1 - sx]);   jv2 = 0.5 * (v[c] + v[c + 1]);
        kv1 = 0.5 * (w[c - sxy] + w[c + 1 - sxy]); kv2 = 0.5 * (w[c] + w[c + 1]);
    }
    else if(ipol == 2)
    {
        iv1 = 0.5 * (u[c - 1] + u[c - 1 + sx]);     iv2 = 0.5 * (u[c] + u[c + sx]);
        jv1 = 0.5 * (v[c] + v[c - sx]);             jv2 = 0.5 * (v[c] + v[c + sx]);
        kv1 = 0.5 * (w[c - sxy] + w[c - sxy + sx]); kv2 = 0.5 * (w[c] + w[c + sx]);
    }
    else
    {
        iv1 = 0.5 * (u[c - 1] + u[c - 1 + sxy]);    iv2 = 0.5 * (u[c] + u[c + sxy]);
        jv1 = 0.5 * (v[c - sx] + v[c - sx + sxy]);  jv2 = 0.5 * (v[c] + v[c + sxy]);
        kv1 = 0.5 * (w[c] + w[c - sxy]);            kv2 = 0.5 * (w[c] + w[c + sxy]);
    }
    const int ufx = ipol == 1, ufy = ipol == 2, ufz = ipol == 3;
    const double* DX = ipol == 1 ? g.DXP : g.DXN;
    const double* DY = ipol == 2 ? g.DYP : g.DYN;
    const double* DZ = ipol == 3 ? g.DZP : g.DZN;
    double fv1 = 0.0, fv2 = 0.0;
    const double fu1 = flux_face(g, 0, ufx, q, c - 1, 1, i - 1, iv1);
    const double fu2 = flux_face(g, 0, ufx, q, c, 1, i, iv2);
    if(g.jdir == 1)
    {
        fv1 = flux_face(g, 1, ufy, q, c - sx, sx, j - 1, jv1);
        fv2 = flux_face(g, 1, ufy, q, c, sx, j, jv2);
    }
    const double fw1 = flux_face(g, 2, ufz, q, c - sxy, sxy, k - 1, kv1);
    const double fw2 = flux_face(g, 2, ufz, q, c, sxy, k, kv2);
    return -((iv2 * fu2 - iv1 * fu1) / DX[i + MARGE]) - ((jv2 * fv2 - jv1 * fv1) / DY[j + MARGE])
         - ((kv2 * fw2 - kv1 * fw1) / DZ[k + MARGE]);
}

__device__ __forceinline__ bool comp_cell(const DevGrid& g, int ipol, int i, int j, int k)
{
    // ULOOP/VLOOP/WLOOP (looping.h:151-153) with the box flags: the last face row of the
    // component is a wall face and is not updated
    if(i >= g.nx || j >= g.ny || k >= g.nz) return false;
    if(ipol == 1 && i >= g.nx - 1) return false;
    if(ipol == 2 && j >= g.ny - 1) return false;
    if(ipol == 3 && k >= g.nz - 1) return false;
    return true;
}

__global__ void __launch_bounds__(256) k_flux_acc(DevGrid g, int nbx, int nby, int nbtot, int ipol,
                                                  const double* __restrict__ q, const double* __restrict__ u,
                                                  const double* __restrict__ v, const double* __restrict__ w,
                                                  double* __restrict__ acc)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(!comp_cell(g, ipol, i, j, k)) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
    acc[c] += flux_div(g, ipol, q, u, v, w, c, i, j, k);
}

void launch_flux_acc(const DevGrid& g, int ipol, const double* q, const double* u, const double* v, const double* w,
                     double* acc, hipStream_t s)
{
    if(ipol == 2 && g.jdir == 0) return;
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_flux_acc, grid, dim3(256), 0, s, g, nbx, nby, nbtot, ipol, q, u, v, w, acc);
}

// density_f::roface (density_f.cpp:35-51)
__device__ __forceinline__ double roface(double phic, double phin, double psi, double rw, double ra)
{
    const double phival = 0.5 * (phic + phin);
    double H = 0.0;
    if(phival > psi) H = 1.0;
    if(phival < -psi) H = 0.0;
    if(fabs(phival) <= psi) H = 0.5 * (1.0 + phival / psi + (1.0 / PI_REF) * sin((PI_REF * phival) / psi));
    return rw * H + ra * (1.0 - H);
}

// Momentum predictor of one velocity component, momentum_FC3.cpp:160-208 (stage 1),
// :262-310 (stage 2), :360-408 (stage 3): F = 0 (kepsilon_void::isource), pressure
// gradient (pjm_corr.cpp:185-204), gravity (irhs, momentum_FC3.cpp:432-472), WENO flux
// (weno_flux_nug), diffusion off (diff_void copies q), then the RK combination.
__global__ void __launch_bounds__(256) k_momentum(DevGrid g, int nbx, int nby, int nbtot, int comp, int stage,
                                                  const double* __restrict__ qin, const double* __restrict__ qold,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, const double* __restrict__ press,
                                                  const double* __restrict__ phi, double grav, double cdt,
                                                  double* __restrict__ out, double rw, double ra)
{
    int bx, by, bz;
    if(!tile_of(nbx, nby, nbtot, bx, by, bz)) return;
    const int i = bx * TX + (threadIdx.x & (TX - 1));
    const int j = by * TY + (threadIdx.x >> 6);
    const int k = bz;
    if(!comp_cell(g, comp, i, j, k)) return;
    const long c = g.base + i + g.sx * j + g.sxy * k;
    const long off = comp == 1 ? 1 : comp == 2 ? g.sx : g.sxy;
    const double* D = comp == 1 ? g.DXP : comp == 2 ? g.DYP : g.DZP;
    const int ix = comp == 1 ? i : comp == 2 ? j : k;
    double F = 0.0;
    F -= 1.0 * (press[c + off] - press[c]) / (D[ix + MARGE] * roface(phi[c], phi[c + off], g.psi, rw, ra));
    F += (0.0 + grav + 0.0 + 0.0) * 1.0;
    const double q = comp == 1 ? u[c] : comp == 2 ? v[c] : w[c];
    (void)q;
    F += flux_div(g, comp, qin, u, v, w, c, i, j, k);
    const double df = qin[c];
    if(stage == 1)
        out[c] = df + cdt * 1.0 * F;
    else if(stage == 2)
        out[c] = 0.75 * qold[c] + 0.25 * df + cdt * 1.0 * F;
    else
        out[c] = (1.0 / 3.0) * qold[c] + (2.0 / 3.0) * df + cdt * 1.0 * F;
}

void launch_momentum(const DevGrid& g, int comp, int stage, const double* qin, const double* qold, const double* u,
                     const double* v, const double* w, const double* press, const double* phi, double grav,
                     double cdt, double* out, const double* rho_w_a, hipStream_t s)
{
    if(comp == 2 && g.jdir == 0) return;
    int nbx, nby, nbtot;
    dim3 grid = tile_grid(g, g.nx, g.ny, g.nz, nbx, nby, nbtot);
    hipLaunchKernelGGL(k_momentum, grid, dim3(256), 0, s, g, nbx, nby, nbtot, comp, stage, qin, qold, u, v, w, press,
                       phi, grav, cdt, out, rho_w_a[0], rho_w_a[1]);
}

}  // namespace reef
