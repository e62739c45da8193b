// Per-cell arithmetic of the hot path, shared by every kernel shell: the WENO5
// reconstruction and smoothness weights, HJ gradients, flux divergence, face density,
// the folded Poisson operator, and the block / finalize reductions.  Each function keeps
// the reference's operation order (the library is built with -ffp-contract=off), so any
// kernel that calls it reproduces the reference's rounding.
#pragma once

#include "reef_internal.h"

namespace reef {

constexpr int TX = 64, TY = 4, NT = 256;

// ------------------------------------------------------------------ WENO5 core
// Smoothness indicators, (IS+psi)^-2 weights without alpha normalisation and the
// non-uniform reconstruction, in the operation order of weno_nug_func.h:48-196 and
// weno_flux_nug.cpp:136-160 / weno_hj_df_nug.cpp:110-177.  The stencil side (minus: the
// upwind stencil from below) is chosen per lane and every lane of a wavefront runs the
// same instructions, so lanes that take different sides do not run the reconstruction
// twice.  Each side's operations and their order are the reference's: the two sides
// differ only in which table entries and differences enter a term, and the reference's
// "x - Q1 (q5 - q4)" is computed as x + (-Q1) (q5 - q4), which IEEE arithmetic rounds
// identically (negation is exact and x - y is x + (-y)).
// Table entry e of I, C, Q is at [e * st]: st = 1 for a cell's own table row (y, z), the
// row length for the entry-major x tables.
__device__ __forceinline__ double weno_recon(const double* __restrict__ I, const double* __restrict__ C,
                                             const double* __restrict__ Q, long st, bool minus, double q1, double q2,
                                             double q3, double q4, double q5)
{
    const double dq12 = q1 - q2, dq23 = q2 - q3, dq32 = q3 - q2, dq34 = q3 - q4, dq43 = q4 - q3, dq54 = q5 - q4;
    const int r = minus ? 0 : 3;
    const double* I1 = I + (long)(r * 3) * st;
    const double is1 = I1[0] * dq54 * dq54 + I1[st] * (dq54) * (dq34) + I1[2 * st] * dq34 * dq34;
    const double is2 = I1[3 * st] * dq23 * dq23 + I1[4 * st] * (dq43) * (dq23) + I1[5 * st] * dq43 * dq43;
    const double is3 = I1[6 * st] * dq12 * dq12 + I1[7 * st] * (dq32) * (dq12) + I1[8 * st] * dq32 * dq32;
    const double a1 = is1 + 1.0e-6, a2 = is2 + 1.0e-6, a3 = is3 + 1.0e-6;
    const double c0 = C[r * st], c1 = C[(r + 1) * st], c2 = C[(r + 2) * st];
    const double w1 = c0 / (0.0 + (a1 * a1) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
    const double w2 = c1 / (0.0 + (a2 * a2) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
    const double w3 = c2 / (0.0 + (a3 * a3) * (c0 / (a1 * a1) + c1 / (a2 * a2) + c2 / (a3 * a3)));
    // minus: w1 (q4 + Q0 (q3 - q4) - Q1 (q5 - q4)) + w2 (q3 + Q2 (q4 - q3) - Q3 (q2 - q3))
    //        + w3 (q2 + Q4 (q1 - q2) + Q5 (q3 - q2))
    // plus:  w1 (q4 + Q6 (q3 - q4) + Q7 (q5 - q4)) + w2 (q3 + Q8 (q2 - q3) - Q9 (q4 - q3))
    //        + w3 (q2 + Q10 (q3 - q2) - Q11 (q1 - q2))
    const double* Qs = Q + (long)(minus ? 0 : 6) * st;
    const double A1 = Qs[0], B1 = minus ? -Qs[st] : Qs[st];
    const double A2 = Qs[2 * st], B2 = Qs[3 * st];
    const double u2 = minus ? dq43 : dq23, v2 = minus ? dq23 : dq43;
    const double A3 = Qs[4 * st], B3 = minus ? Qs[5 * st] : -Qs[5 * st];
    const double u3 = minus ? dq12 : dq32, v3 = minus ? dq32 : dq12;
    return w1 * (q4 + A1 * dq34 + B1 * dq54) + w2 * (q3 + A2 * u2 - B2 * v2) + w3 * (q2 + A3 * u3 + B3 * v3);
}

// HJ gradient along one axis at index ix (stride s) on the side `minus` chosen per lane,
// given the two central divided differences dd2 = (f0 - fm1) / D(ip-1) and
// dd3 = (fp1 - f0) / D(ip) (weno_hj_df_nug.cpp:344-396 / ddweno_nug_sf.cpp:186-224): the
// minus stencil takes (fm2 - fm3) / D(ip-3), dd1, dd2, dd3, dd4, the plus stencil dd1, dd2,
// dd3, dd4, (fp3 - fp2) / D(ip+2); only the outer quotient of the lane's side is formed
__device__ __forceinline__ double hj_grad_dd(const DevGrid& g, int axis, const double* __restrict__ f, long c, long s,
                                             int ix, bool minus, double dd2, double dd3)
{
    const int ip = ix + MARGE;
    const double* D = axis == 0 ? g.DXP : axis == 1 ? g.DYP : g.DZP;
    const long st = axis == 0 ? g.nT : 1;
    const double* I = axis == 0 ? g.isfT + ip : g.isf[axis] + (size_t)ip * 36;
    const double* C = axis == 0 ? g.cfT + ip : g.cf[axis] + (size_t)ip * 12;
    const double* Q = axis == 0 ? g.qfT + ip : g.qf[axis] + (size_t)ip * 24;
    const double fm2 = f[c - 2 * s], fm1 = f[c - s], fp1 = f[c + s], fp2 = f[c + 2 * s];
    const double fo = f[minus ? c - 3 * s : c + 3 * s];
    const double dd1 = (fm1 - fm2) / D[ip - 2];
    const double dd4 = (fp2 - fp1) / D[ip + 1];
    const double ddo = (minus ? fm2 - fo : fo - fp2) / D[minus ? ip - 3 : ip + 2];
    return weno_recon(I, C, Q, st, minus, minus ? ddo : dd1, minus ? dd1 : dd2, minus ? dd2 : dd3, minus ? dd3 : dd4,
                      minus ? dd4 : ddo);
}

// the HJ gradient upwind of advec (minus for advec > 0, plus for advec < 0, 0 otherwise)
__device__ __forceinline__ double hj_grad(const DevGrid& g, int axis, const double* __restrict__ f, long c, long s,
                                          int ix, double advec)
{
    const bool minus = advec > 0.0;
    if(!minus && !(advec < 0.0)) return 0.0;
    const int ip = ix + MARGE;
    const double* D = axis == 0 ? g.DXP : axis == 1 ? g.DYP : g.DZP;
    const double f0 = f[c];
    return hj_grad_dd(g, axis, f, c, s, ix, minus, (f0 - f[c - s]) / D[ip - 1], (f[c + s] - f0) / D[ip]);
}

// ------------------------------------------------------------------ WENO flux
// weno_flux_nug::fx/fy/fz with iqmin..kqmax (weno_flux_nug.cpp:136-313): values; the
// x upwind-from-above stencil repeats f(i+2) (weno_flux_nug.cpp:288-295).  The side is
// chosen per lane (weno_recon).
__device__ __forceinline__ double flux_face(const DevGrid& g, int axis, int uf, const double* __restrict__ f, long c,
                                            long s, int ix, double advec)
{
    const bool minus = advec > 0.0;
    if(!minus && !(advec < 0.0)) return 0.0;
    const int ip = ix + MARGE;
    const long st = axis == 0 ? g.nT : 1;
    const double* I = axis == 0 ? g.isfT + (long)(uf * 18) * st + ip : g.isf[axis] + (size_t)ip * 36 + uf * 18;
    const double* C = axis == 0 ? g.cfT + (long)(uf * 6) * st + ip : g.cf[axis] + (size_t)ip * 12 + uf * 6;
    const double* Q = axis == 0 ? g.qfT + (long)(uf * 12) * st + ip : g.qf[axis] + (size_t)ip * 24 + uf * 12;
    const double fm1 = f[c - s], f0 = f[c], fp1 = f[c + s], fp2 = f[c + 2 * s];
    const double fo = f[minus ? c - 2 * s : (axis == 0 ? c + 2 * s : c + 3 * s)];
    return weno_recon(I, C, Q, st, minus, minus ? fo : fm1, minus ? fm1 : f0, minus ? f0 : fp1, minus ? fp1 : fp2,
                      minus ? fp2 : fo);
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

// Block partial sums of M values in a fixed tree order, written to
// partials[blockIdx.x*M + m].  The totals are formed by a one-block finalize kernel
// launched right after (sum_partials): a kernel boundary is a far cheaper way to make
// the partials visible than a device-scope fence per block (each such fence writes the
// L2 back on this GPU).
template<int M>
__device__ __forceinline__ void block_partials(const double (&v)[M], double* __restrict__ partials)
{
    __shared__ double sh[M][NT];
    const int t = threadIdx.x;
    for(int m = 0; m < M; ++m) sh[m][t] = v[m];
    __syncthreads();
    for(int s = NT / 2; s > 0; s >>= 1)
    {
        if(t < s)
            for(int m = 0; m < M; ++m) sh[m][t] += sh[m][t + s];
        __syncthreads();
    }
    if(t == 0)
        for(int m = 0; m < M; ++m) partials[(size_t)blockIdx.x * M + m] = sh[m][0];
}

// The same partial sums with a wavefront butterfly (fixed order, lane 0's sum) and a
// 4-entry LDS combine: needs M * 4 doubles of LDS instead of M * NT, so kernels that keep
// stencil planes in LDS still fit eight blocks per CU.
template<int M>
__device__ __forceinline__ void block_partials_w(double (&v)[M], double* __restrict__ partials)
{
    __shared__ double ws[M][NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(int m = 0; m < M; ++m)
        for(int off = 32; off > 0; off >>= 1) v[m] += __shfl_xor(v[m], off);
    if(lane == 0)
        for(int m = 0; m < M; ++m) ws[m][wave] = v[m];
    __syncthreads();
    if(threadIdx.x == 0)
        for(int m = 0; m < M; ++m)
            partials[(size_t)blockIdx.x * M + m] = ((ws[m][0] + ws[m][1]) + ws[m][2]) + ws[m][3];
}

// totals of nb block partials by one block of NT threads: each thread adds a strided
// subset in a fixed order, the wavefronts combine by a butterfly and the four wave sums
// in a fixed order, so every thread gets the same totals with one block barrier
template<int M>
__device__ __forceinline__ void sum_partials(const double* __restrict__ partials, int nb, double (&tot)[M])
{
    __shared__ double ws[M][NT / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double acc[M];
    for(int m = 0; m < M; ++m) acc[m] = 0.0;
    int b = t;
    for(; b + 3 * NT < nb; b += 4 * NT)
    {
        double q[4][M];
        for(int u = 0; u < 4; ++u)
            for(int m = 0; m < M; ++m) q[u][m] = partials[(size_t)(b + u * NT) * M + m];
        for(int u = 0; u < 4; ++u)
            for(int m = 0; m < M; ++m) acc[m] += q[u][m];
    }
    for(; b < nb; b += NT)
        for(int m = 0; m < M; ++m) acc[m] += partials[(size_t)b * M + m];
    for(int m = 0; m < M; ++m)
        for(int off = 32; off > 0; off >>= 1) acc[m] += __shfl_xor(acc[m], off);
    if(lane == 0)
        for(int m = 0; m < M; ++m) ws[m][wave] = acc[m];
    __syncthreads();
    for(int m = 0; m < M; ++m) tot[m] = ((ws[m][0] + ws[m][1]) + ws[m][2]) + ws[m][3];
}

// Couplings of poisson_pcorr.cpp:95-123 before the boundary folding: T[0..5] for the
// +x, -x, +y, -y, +z, -z faces; face densities density_f::roface stored per "+" face.
template<class V>
struct CouplingV
{
    V T[6];
};
using Coupling = CouplingV<double>;

__device__ __forceinline__ Coupling couplings(const DevGrid& g, const double* __restrict__ rfx,
                                              const double* __restrict__ rfy, const double* __restrict__ rfz, long c,
                                              int i, int j, int k)
{
    const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
    Coupling q;
    q.T[0] = (1.0 * 1.0) / (rfx[c] * g.DXP[ip] * g.DXN[ip]);
    q.T[1] = (1.0 * 1.0) / (rfx[c - 1] * g.DXP[ip - 1] * g.DXN[ip]);
    q.T[2] = (1.0 * 1.0) / (rfy[c] * g.DYP[jp] * g.DYN[jp]);
    q.T[3] = (1.0 * 1.0) / (rfy[c - g.sx] * g.DYP[jp - 1] * g.DYN[jp]);
    q.T[4] = (1.0 * 1.0) / (rfz[c] * g.DZP[kp] * g.DZN[kp]);
    q.T[5] = (1.0 * 1.0) / (rfz[c - g.sxy] * g.DZP[kp - 1] * g.DZN[kp]);
    return q;
}

// The FP64 operator recomputes its couplings from the face densities per application
// (bit-identical to the reference's assembly).  The FP32 operator of the mixed-precision
// inner solve reads face conductances 1/(rho_f * DP) rounded to FP32 and scales them by
// the FP32 1/DN of the cell: T[0] = kx(c) / DXN(i), T[1] = kx(c-1) / DXN(i), ...
__device__ __forceinline__ CouplingV<double> coupling_of(const DevGrid& g, const SolverOp64& op, long c, int i, int j,
                                                         int k)
{
    return couplings(g, op.rfx, op.rfy, op.rfz, c, i, j, k);
}

__device__ __forceinline__ CouplingV<float> coupling_of(const DevGrid& g, const SolverOp32& op, long c, int i, int j,
                                                        int k)
{
    const float* __restrict__ kx = op.kx;
    const float* __restrict__ ky = op.ky;
    const float* __restrict__ kz = op.kz;
    const float ix = op.iDXN[i + MARGE], iy = op.iDYN[j + MARGE], iz = op.iDZN[k + MARGE];
    CouplingV<float> q;
    q.T[0] = kx[c] * ix;
    q.T[1] = kx[c - 1] * ix;
    q.T[2] = ky[c] * iy;
    q.T[3] = ky[c - g.sx] * iy;
    q.T[4] = kz[c] * iz;
    q.T[5] = kz[c - g.sxy] * iz;
    return q;
}

// The grid factors of the couplings along one axis at index ix: DP(ix), DN(ix), DP(ix-1)
// for the FP64 operator, the FP32 1/DN(ix) for the FP32 one.  The plane walk loads the
// x and y factors once per column and the z factors with each plane.
template<class Op>
struct AxisFactor;
template<>
struct AxisFactor<SolverOp64>
{
    double p0, n, p1;
};
template<>
struct AxisFactor<SolverOp32>
{
    float inv;
};

__device__ __forceinline__ AxisFactor<SolverOp64> axis_factor(const DevGrid& g, const SolverOp64&, int axis, int ix)
{
    const double* P = axis == 0 ? g.DXP : axis == 1 ? g.DYP : g.DZP;
    const double* N = axis == 0 ? g.DXN : axis == 1 ? g.DYN : g.DZN;
    const int ip = ix + MARGE;
    return AxisFactor<SolverOp64>{P[ip], N[ip], P[ip - 1]};
}

__device__ __forceinline__ AxisFactor<SolverOp32> axis_factor(const DevGrid&, const SolverOp32& op, int axis, int ix)
{
    const float* I = axis == 0 ? op.iDXN : axis == 1 ? op.iDYN : op.iDZN;
    return AxisFactor<SolverOp32>{I[ix + MARGE]};
}

// couplings from the six face values (+x, -x, +y, -y, +z, -z) and the axis factors of the
// cell: the face densities for the FP64 operator (the operation order of couplings()),
// the FP32 conductances for the FP32 one (that of coupling_of)
__device__ __forceinline__ CouplingV<double> coupling_axes(const SolverOp64&, double F0, double F1, double F2,
                                                           double F3, double F4, double F5,
                                                           const AxisFactor<SolverOp64>& ax,
                                                           const AxisFactor<SolverOp64>& ay,
                                                           const AxisFactor<SolverOp64>& az)
{
    CouplingV<double> q;
    q.T[0] = (1.0 * 1.0) / (F0 * ax.p0 * ax.n);
    q.T[1] = (1.0 * 1.0) / (F1 * ax.p1 * ax.n);
    q.T[2] = (1.0 * 1.0) / (F2 * ay.p0 * ay.n);
    q.T[3] = (1.0 * 1.0) / (F3 * ay.p1 * ay.n);
    q.T[4] = (1.0 * 1.0) / (F4 * az.p0 * az.n);
    q.T[5] = (1.0 * 1.0) / (F5 * az.p1 * az.n);
    return q;
}

__device__ __forceinline__ CouplingV<float> coupling_axes(const SolverOp32&, float F0, float F1, float F2, float F3,
                                                          float F4, float F5, const AxisFactor<SolverOp32>& ax,
                                                          const AxisFactor<SolverOp32>& ay,
                                                          const AxisFactor<SolverOp32>& az)
{
    CouplingV<float> q;
    q.T[0] = F0 * ax.inv;
    q.T[1] = F1 * ax.inv;
    q.T[2] = F2 * ay.inv;
    q.T[3] = F3 * ay.inv;
    q.T[4] = F4 * az.inv;
    q.T[5] = F5 * az.inv;
    return q;
}

// the x and z couplings of a 2-D (x-z) cell; the y couplings are 0 (the operator weights
// them by 0)
__device__ __forceinline__ CouplingV<double> coupling_xz(const SolverOp64&, double F0, double F1, double F4,
                                                         double F5, const AxisFactor<SolverOp64>& ax,
                                                         const AxisFactor<SolverOp64>& az)
{
    CouplingV<double> q;
    q.T[0] = (1.0 * 1.0) / (F0 * ax.p0 * ax.n);
    q.T[1] = (1.0 * 1.0) / (F1 * ax.p1 * ax.n);
    q.T[2] = 0.0;
    q.T[3] = 0.0;
    q.T[4] = (1.0 * 1.0) / (F4 * az.p0 * az.n);
    q.T[5] = (1.0 * 1.0) / (F5 * az.p1 * az.n);
    return q;
}

__device__ __forceinline__ CouplingV<float> coupling_xz(const SolverOp32&, float F0, float F1, float F4, float F5,
                                                        const AxisFactor<SolverOp32>& ax,
                                                        const AxisFactor<SolverOp32>& az)
{
    CouplingV<float> q;
    q.T[0] = F0 * ax.inv;
    q.T[1] = F1 * ax.inv;
    q.T[2] = 0.0f;
    q.T[3] = 0.0f;
    q.T[4] = F4 * az.inv;
    q.T[5] = F5 * az.inv;
    return q;
}

__device__ __forceinline__ const double* face_x(const SolverOp64& op) { return op.rfx; }
__device__ __forceinline__ const double* face_y(const SolverOp64& op) { return op.rfy; }
__device__ __forceinline__ const double* face_z(const SolverOp64& op) { return op.rfz; }
__device__ __forceinline__ const float* face_x(const SolverOp32& op) { return op.kx; }
__device__ __forceinline__ const float* face_y(const SolverOp32& op) { return op.ky; }
__device__ __forceinline__ const float* face_z(const SolverOp32& op) { return op.kz; }

template<class V>
__device__ __forceinline__ V diag_of(const DevGrid& g, const CouplingV<V>& q)
{
    const V yd = (V)g.ydir;
    return q.T[0] + q.T[1] + q.T[2] * yd + q.T[3] * yd + q.T[4] + q.T[5];
}

// y = A x at cell c with the folded operator of poisson_pcorr.cpp:124-230: boundary
// neighbours (pressure-correction ghosts are 0) drop out, the diagonal keeps them.
// Summation order of matvec_std (bicgstab_ijk.cpp:296-310); 5 points in 2-D
// (bicgstab_ijk_2D.cpp:308-320).  xv(cc) gives the value of x at element cc.
// the same sum from the seven values (xn/xs: i+-1, xw/xe: j+-1, xt/xb: k+-1), values
// outside the domain are ignored
template<class V>
__device__ __forceinline__ V stencil7(const DevGrid& g, const CouplingV<V>& q, V xc, V xn_, V xs_, V xw_, V xe_, V xt_,
                                      V xb_, int i, int j, int k)
{
    const V Mp = diag_of(g, q);
    const V Mn = i == g.nx - 1 ? V(0) : -q.T[0];
    const V Ms = i == 0 ? V(0) : -q.T[1];
    const V Mt = k == g.nz - 1 ? V(0) : -q.T[4];
    const V Mb = k == 0 ? V(0) : -q.T[5];
    const V xn = i == g.nx - 1 ? V(0) : xn_;
    const V xs = i == 0 ? V(0) : xs_;
    const V xt = k == g.nz - 1 ? V(0) : xt_;
    const V xb = k == 0 ? V(0) : xb_;
    const V yd = (V)g.ydir;
    const V Mw = j == g.ny - 1 ? V(0) : -q.T[2] * yd;
    const V Me = j == 0 ? V(0) : -q.T[3] * yd;
    const V xw = j == g.ny - 1 ? V(0) : xw_;
    const V xe = j == 0 ? V(0) : xe_;
    return Mp * xc + Mn * xn + Ms * xs + Mw * xw + Me * xe + Mt * xt + Mb * xb;
}

// the 2-D (x-z) operator of apply_A_fn from the five values (xn/xs: i+-1, xt/xb: k+-1)
template<class V>
__device__ __forceinline__ V stencil5(const DevGrid& g, const CouplingV<V>& q, V xc, V xn_, V xs_, V xt_, V xb_, int i,
                                      int k)
{
    const V Mp = diag_of(g, q);
    const V Mn = i == g.nx - 1 ? V(0) : -q.T[0];
    const V Ms = i == 0 ? V(0) : -q.T[1];
    const V Mt = k == g.nz - 1 ? V(0) : -q.T[4];
    const V Mb = k == 0 ? V(0) : -q.T[5];
    const V xn = i == g.nx - 1 ? V(0) : xn_;
    const V xs = i == 0 ? V(0) : xs_;
    const V xt = k == g.nz - 1 ? V(0) : xt_;
    const V xb = k == 0 ? V(0) : xb_;
    return Mp * xc + Mn * xn + Ms * xs + Mt * xt + Mb * xb;
}

template<class V, class XV>
__device__ __forceinline__ V apply_A_fn(const DevGrid& g, const CouplingV<V>& q, XV xv, long c, int i, int j, int k)
{
    const V Mp = diag_of(g, q);
    const V Mn = i == g.nx - 1 ? V(0) : -q.T[0];
    const V Ms = i == 0 ? V(0) : -q.T[1];
    const V Mt = k == g.nz - 1 ? V(0) : -q.T[4];
    const V Mb = k == 0 ? V(0) : -q.T[5];
    const V xn = i == g.nx - 1 ? V(0) : xv(c + 1);
    const V xs = i == 0 ? V(0) : xv(c - 1);
    const V xt = k == g.nz - 1 ? V(0) : xv(c + g.sxy);
    const V xb = k == 0 ? V(0) : xv(c - g.sxy);
    if(g.jdir == 1)
    {
        const V yd = (V)g.ydir;
        const V Mw = j == g.ny - 1 ? V(0) : -q.T[2] * yd;
        const V Me = j == 0 ? V(0) : -q.T[3] * yd;
        const V xw = j == g.ny - 1 ? V(0) : xv(c + g.sx);
        const V xe = j == 0 ? V(0) : xv(c - g.sx);
        return Mp * xv(c) + Mn * xn + Ms * xs + Mw * xw + Me * xe + Mt * xt + Mb * xb;
    }
    return Mp * xv(c) + Mn * xn + Ms * xs + Mt * xt + Mb * xb;
}

template<class V>
__device__ __forceinline__ V apply_A(const DevGrid& g, const CouplingV<V>& q, const V* __restrict__ x, long c, int i,
                                     int j, int k)
{
    return apply_A_fn(g, q, [&](long cc) { return x[cc]; }, c, i, j, k);
}


// HJ-WENO convection term of phi at a cell: weno_hj_df_nug::aij (weno_hj_df_nug.cpp:85-108)
// with the cell-centred flux_HJ_CDS2 velocities (ipol 4, flux_HJ_CDS2.cpp:59-127)
__device__ __forceinline__ double hj_L(const DevGrid& g, const double* __restrict__ phi_s, const double* __restrict__ u,
                                       const double* __restrict__ v, const double* __restrict__ w, long c, int i, int j,
                                       int k)
{
    const double iadvec = 0.5 * (u[c] + u[c - 1]);
    const double jadvec = 0.5 * (v[c] + v[c - g.sx]);
    const double kadvec = 0.5 * (w[c] + w[c - g.sxy]);
    double L = -iadvec * hj_grad(g, 0, phi_s, c, 1, i, iadvec);
    if(g.jdir == 1) L -= jadvec * hj_grad(g, 1, phi_s, c, g.sx, j, jadvec);
    L -= kadvec * hj_grad(g, 2, phi_s, c, g.sxy, k, kadvec);
    return L;
}

// reinidisc_f::disc (reinidisc_f.cpp:57-127) at a cell; fsf = 1 is reinidisc_fsf::disc
// (reinidisc_fsf.cpp:66-141, F49 = 0): the same update on cells whose 7-point stencil has
// one sign, L = 0 elsewhere (ddweno_nug and ddweno_nug_sf in modus 0 compute the same
// gradients)
__device__ __forceinline__ double reini_L(const DevGrid& g, const double* __restrict__ f, long c, int i, int j, int k,
                                          int fsf)
{
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
        // the side of each axis (reinidisc_f.cpp:78-104): minus, plus or none (the three tests
        // exclude each other but for the last, which zeroes the gradient); the quotients of
        // the tests are the central divided differences of the WENO stencil
        auto axis_grad = [&](int axis, long st, int ix, const double* DP, int pp) {
            const double dmin = (lsv - f[c - st]) / DP[pp - 1];
            const double dplus = (f[c + st] - lsv) / DP[pp];
            const bool cm = dmin * lsSig > 0.0 && dplus * lsSig > -dmin * lsSig;
            const bool cp = dplus * lsSig < 0.0 && dmin * lsSig < -dplus * lsSig;
            const bool cz = dplus * lsSig > 0.0 && dmin * lsSig < 0.0;
            double d = 0.0;
            if((cm || cp) && !cz) d = hj_grad_dd(g, axis, f, c, st, ix, cm, dmin, dplus);
            return d;
        };
        dx = axis_grad(0, 1, i, g.DXP, ip);
        if(g.jdir == 1) dy = axis_grad(1, g.sx, j, g.DYP, jp);
        dz = axis_grad(2, g.sxy, k, g.DZP, kp);
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
    return L;
}

// per-cell pseudo time step F43 * min(DXP, DYP, DZP) (reini_RK3.cpp:150-160)
__device__ __forceinline__ double reini_dtc(const DevGrid& g, int i, int j, int k, double F43)
{
    const double a = g.DXP[i + MARGE], b = g.DYP[j + MARGE], cc = g.DZP[k + MARGE];
    const double m = g.jdir == 0 ? (a < cc ? a : cc) : (((a < b ? a : b) < cc) ? (a < b ? a : b) : cc);
    return F43 * m;
}

// right-hand side of one velocity component in momentum_FC3 (momentum_FC3.cpp:160-208 and
// the later stages): F = 0 (kepsilon_void::isource), pressure gradient
// (pjm_corr.cpp:185-204), gravity (irhs, momentum_FC3.cpp:432-472), WENO flux divergence
__device__ __forceinline__ double momentum_F(const DevGrid& g, int comp, const double* __restrict__ qin,
                                             const double* __restrict__ u, const double* __restrict__ v,
                                             const double* __restrict__ w, const double* __restrict__ press,
                                             const double* __restrict__ phi, double grav, double rw, double ra, long c,
                                             int i, int j, int k)
{
    const long off = comp == 1 ? 1 : comp == 2 ? g.sx : g.sxy;
    const double* D = comp == 1 ? g.DXP : comp == 2 ? g.DYP : g.DZP;
    const int ix = comp == 1 ? i : comp == 2 ? j : k;
    double F = 0.0;
    F -= 1.0 * (press[c + off] - press[c]) / (D[ix + MARGE] * roface(phi[c], phi[c + off], g.psi, rw, ra));
    F += (0.0 + grav + 0.0 + 0.0) * 1.0;
    F += flux_div(g, comp, qin, u, v, w, c, i, j, k);
    return F;
}

// ediff2::diff_u / diff_v / diff_w (ediff2.cpp:78-190): the explicit viscous term of
// component comp at cell c, which momentum_FC3 adds to F after the convection (D20 = 1).
// Laminar: eddyv = 0, and visc + 0.0 is visc exactly, so the reference's visc + eddyv sums
// are the visc sums in the same order.
__device__ __forceinline__ double ediff_term(const DevGrid& g, int comp, const double* __restrict__ u,
                                             const double* __restrict__ v, const double* __restrict__ w,
                                             const double* __restrict__ nu, long c, int i, int j, int k)
{
    const long X = 1, Y = g.sx, Z = g.sxy;
    const int ip = i + MARGE, jp = j + MARGE, kp = k + MARGE;
    auto V = [&](long o) { return nu[c + o]; };
    if(comp == 1)
    {
        const double uc = u[c];
        const double ddy_p = (((V(0) + V(X)) + V(Y)) + V(X + Y)) * 0.25;
        const double ddy_m = (((V(-Y) + V(X - Y)) + V(0)) + V(X)) * 0.25;
        const double ddz_p = (((V(0) + V(X)) + V(Z)) + V(X + Z)) * 0.25;
        const double ddz_m = (((V(-Z) + V(X - Z)) + V(0)) + V(X)) * 0.25;
        return 2.0 * ((u[c + X] - uc) * (V(X) / g.DXN[ip + 1]) - (uc - u[c - X]) * (V(0) / g.DXN[ip])) / g.DXP[ip]
               + ((u[c + Y] - uc) * (ddy_p / g.DYP[jp]) - (uc - u[c - Y]) * (ddy_m / g.DYP[jp - 1])) / g.DYN[jp]
               + ((u[c + Z] - uc) * (ddz_p / g.DZP[kp]) - (uc - u[c - Z]) * (ddz_m / g.DZP[kp - 1])) / g.DZN[kp]
               + ((v[c + X] - v[c]) * (ddy_p / g.DXP[ip]) - (v[c + X - Y] - v[c - Y]) * (ddy_m / g.DXP[ip])) / g.DYN[jp]
               + ((w[c + X] - w[c]) * (ddz_p / g.DXP[ip]) - (w[c + X - Z] - w[c - Z]) * (ddz_m / g.DXP[ip])) / g.DZN[kp];
    }
    if(comp == 2)
    {
        const double vc = v[c];
        const double ddx_p = (((V(0) + V(Y)) + V(X)) + V(X + Y)) * 0.25;
        const double ddx_m = (((V(-X) + V(Y - X)) + V(0)) + V(Y)) * 0.25;
        const double ddz_p = (((V(0) + V(Y)) + V(Z)) + V(Y + Z)) * 0.25;
        const double ddz_m = (((V(-Z) + V(Y - Z)) + V(0)) + V(Y)) * 0.25;
        return ((v[c + X] - vc) * (ddx_p / g.DXP[ip]) - (vc - v[c - X]) * (ddx_m / g.DXP[ip - 1])) / g.DXN[ip]
               + 2.0 * ((v[c + Y] - vc) * (V(Y) / g.DYN[jp + 1]) - (vc - v[c - Y]) * (V(0) / g.DYN[jp])) / g.DYP[jp]
               + ((v[c + Z] - vc) * (ddz_p / g.DZP[kp]) - (vc - v[c - Z]) * (ddz_m / g.DZP[kp - 1])) / g.DZN[kp]
               + ((u[c + Y] - u[c]) * (ddx_p / g.DYP[jp]) - (u[c - X + Y] - u[c - X]) * (ddx_m / g.DYP[jp])) / g.DXN[ip]
               + ((w[c + Y] - w[c]) * (ddz_p / g.DYP[jp]) - (w[c + Y - Z] - w[c - Z]) * (ddz_m / g.DYP[jp])) / g.DZN[kp];
    }
    const double wc = w[c];
    const double ddx_p = (((V(0) + V(Z)) + V(X)) + V(X + Z)) * 0.25;
    const double ddx_m = (((V(-X) + V(Z - X)) + V(0)) + V(Z)) * 0.25;
    const double ddy_p = (((V(0) + V(Z)) + V(Y)) + V(Y + Z)) * 0.25;
    const double ddy_m = (((V(-Y) + V(Z - Y)) + V(0)) + V(Z)) * 0.25;
    return ((w[c + X] - wc) * (ddx_p / g.DXP[ip]) - (wc - w[c - X]) * (ddx_m / g.DXP[ip - 1])) / g.DXN[ip]
           + ((w[c + Y] - wc) * (ddy_p / g.DYP[jp]) - (wc - w[c - Y]) * (ddy_m / g.DYP[jp - 1])) / g.DYN[jp]
           + 2.0 * ((w[c + Z] - wc) * (V(Z) / g.DZN[kp + 1]) - (wc - w[c - Z]) * (V(0) / g.DZN[kp])) / g.DZP[kp]
           + ((u[c + Z] - u[c]) * (ddx_p / g.DZP[kp]) - (u[c - X + Z] - u[c - X]) * (ddx_m / g.DZP[kp])) / g.DXN[ip]
           + ((v[c + Z] - v[c]) * (ddy_p / g.DZP[kp]) - (v[c - Y + Z] - v[c - Y]) * (ddy_m / g.DZP[kp])) / g.DYN[jp];
}

// smoothed Heaviside of fluid_update_fsf (fluid_update_fsf.cpp:52-62)
__device__ __forceinline__ double heaviside_H(double p, double psi)
{
    double H = 0.0;
    if(p > psi) H = 1.0;
    if(p < -psi) H = 0.0;
    if(fabs(p) <= psi) H = 0.5 * (1.0 + p / psi + (1.0 / PI_REF) * sin((PI_REF * p) / psi));
    return H;
}

// pjm_corr::rhs (pjm_corr.cpp:151-176): -div(u*)/(alpha dt)
__device__ __forceinline__ double poisson_rhs_cell(const DevGrid& g, const double* __restrict__ u,
                                                   const double* __restrict__ v, const double* __restrict__ w,
                                                   double adt, long c, int i, int j, int k)
{
    return -(u[c] - u[c - 1]) / (adt * g.DXN[i + MARGE]) - (v[c] - v[c - g.sx]) / (adt * g.DYN[j + MARGE]) * g.ydir
         - (w[c] - w[c - g.sxy]) / (adt * g.DZN[k + MARGE]);
}

// pjm_corr::presscorr + ucorr/vcorr/wcorr (pjm_corr.cpp:123-149) at a cell
__device__ __forceinline__ void correct_cell(const DevGrid& g, const double* __restrict__ pcorr,
                                             const double* __restrict__ phi, double rw, double ra, double adt,
                                             double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                                             double* __restrict__ press, long c, int i, int j, int k)
{
    press[c] += pcorr[c];
    if(i < g.nx - 1)
        u[c] -= adt * 1.0 * 1.0 * ((pcorr[c + 1] - pcorr[c]) / (g.DXP[i + MARGE] * roface(phi[c], phi[c + 1], g.psi, rw, ra)));
    if(g.jdir == 1 && j < g.ny - 1)
        v[c] -= adt * 1.0 * 1.0
              * ((pcorr[c + g.sx] - pcorr[c]) / (g.DYP[j + MARGE] * roface(phi[c], phi[c + g.sx], g.psi, rw, ra)));
    if(k < g.nz - 1)
        w[c] -= adt * 1.0 * 1.0
              * ((pcorr[c + g.sxy] - pcorr[c]) / (g.DZP[k + MARGE] * roface(phi[c], phi[c + g.sxy], g.psi, rw, ra)));
}


// One row of the implicit viscous operator of idiff2_FS::diff_u / diff_v / diff_w
// (idiff2_FSu.cpp:40-170, idiff2_FSv.cpp:29-160, idiff2_FSw.cpp:29-160) and of idiff2_FS_2D
// diff_u / diff_w (idiff2_FSu_2D.cpp, idiff2_FSw_2D.cpp) at cell c of component comp's loop,
// in the reference's operation order.  Laminar (eddyv = 0: each visc + eddyv sum is the visc
// sum), no porosity (CPOR = 1); rhs starts from the 0 that irhs left.  D22 = 1: once the
// direct-forcing flags are set (df), a coupling to a neighbour outside the component's grid
// (outside [0, ie) x [0, je) x [0, ke)) moves to the right-hand side times the cell's own
// value, in the order s, n, e, w, b, t.  diff_w reads the fdm's u and v (au, av) at k + 1.
struct MatRow
{
    double p, n, s, w, e, t, b, rhs;
};

__device__ __forceinline__ MatRow idiff_row(const DevGrid& g, int comp, double alpha, double dt,
                                            const double* __restrict__ vin, const double* __restrict__ u,
                                            const double* __restrict__ v, const double* __restrict__ w,
                                            const double* __restrict__ au, const double* __restrict__ av,
                                            const double* __restrict__ nu, bool df, int ie, int je, int ke, long c, int i,
                                            int j, int k)
{
    const long X = 1, Y = g.sx, Z = g.sxy;
    const int IP = i + MARGE, JP = j + MARGE, KP = k + MARGE;
    const double *DXN = g.DXN, *DYN = g.DYN, *DZN = g.DZN, *DXP = g.DXP, *DYP = g.DYP, *DZP = g.DZP;
    const double CPOR = 1.0;
    const bool two = g.jdir == 0;
    auto V = [&](long o) { return nu[c + o]; };
    MatRow r;
    r.w = r.e = 0.0;
    if(comp == 1)
    {
        const double ddz_p = 0.25 * (((V(0) + V(X)) + V(Z)) + V(X + Z));
        const double ddz_m = 0.25 * (((V(-Z) + V(X - Z)) + V(0)) + V(X));
        if(!two)
        {
            const double ddy_p = 0.25 * (((V(0) + V(X)) + V(Y)) + V(X + Y));
            const double ddy_m = 0.25 * (((V(-Y) + V(X - Y)) + V(0)) + V(X));
            r.p = 2.0 * V(X) / (DXN[IP + 1] * DXP[IP]) + 2.0 * V(0) / (DXN[IP] * DXP[IP]) + ddy_p / (DYP[JP] * DYN[JP])
                + ddy_m / (DYP[JP - 1] * DYN[JP]) + ddz_p / (DZP[KP] * DZN[KP]) + ddz_m / (DZP[KP - 1] * DZN[KP])
                + CPOR / (alpha * dt);
            r.rhs = 0.0 + (((v[c + X] - v[c]) * ddy_p - (v[c + X - Y] - v[c - Y]) * ddy_m) / (DXP[IP] * DYN[JP])
                           + ((w[c + X] - w[c]) * ddz_p - (w[c + X - Z] - w[c - Z]) * ddz_m) / (DXP[IP] * DZN[KP])
                           + (CPOR * vin[c]) / (alpha * dt));
            r.e = -ddy_m / (DYP[JP - 1] * DYN[JP]);
            r.w = -ddy_p / (DYP[JP] * DYN[JP]);
        }
        else
        {
            r.p = 2.0 * V(X) / (DXN[IP + 1] * DXP[IP]) + 2.0 * V(0) / (DXN[IP] * DXP[IP]) + ddz_p / (DZP[KP] * DZN[KP])
                + ddz_m / (DZP[KP - 1] * DZN[KP]) + CPOR / (alpha * dt);
            r.rhs = 0.0 + (((w[c + X] - w[c]) * ddz_p - (w[c + X - Z] - w[c - Z]) * ddz_m) / (DXP[IP] * DZN[KP])
                           + (CPOR * vin[c]) / (alpha * dt));
        }
        r.s = -2.0 * V(0) / (DXN[IP] * DXP[IP]);
        r.n = -2.0 * V(X) / (DXN[IP + 1] * DXP[IP]);
        r.b = -ddz_m / (DZP[KP - 1] * DZN[KP]);
        r.t = -ddz_p / (DZP[KP] * DZN[KP]);
    }
    else if(comp == 2)
    {
        const double ddx_p = 0.25 * (((V(0) + V(Y)) + V(X)) + V(X + Y));
        const double ddx_m = 0.25 * (((V(-X) + V(Y - X)) + V(0)) + V(Y));
        const double ddz_p = 0.25 * (((V(0) + V(Y)) + V(Z)) + V(Y + Z));
        const double ddz_m = 0.25 * (((V(-Z) + V(Y - Z)) + V(0)) + V(Y));
        r.p = 2.0 * V(Y) / (DYN[JP + 1] * DYP[JP]) + 2.0 * V(0) / (DYN[JP] * DYP[JP]) + ddx_p / (DXP[IP] * DXN[IP])
            + ddx_m / (DXP[IP - 1] * DXN[IP]) + ddz_p / (DZP[KP] * DZN[KP]) + ddz_m / (DZP[KP - 1] * DZN[KP])
            + CPOR / (alpha * dt);
        r.rhs = 0.0 + (((u[c + Y] - u[c]) * ddx_p - (u[c - X + Y] - u[c - X]) * ddx_m) / (DYP[JP] * DXN[IP])
                       + ((w[c + Y] - w[c]) * ddz_p - (w[c + Y - Z] - w[c - Z]) * ddz_m) / (DYP[JP] * DZN[KP])
                       + (CPOR * vin[c]) / (alpha * dt));
        r.s = -ddx_m / (DXP[IP - 1] * DXN[IP]);
        r.n = -ddx_p / (DXP[IP] * DXN[IP]);
        r.e = -2.0 * V(0) / (DYN[JP] * DYP[JP]);
        r.w = -2.0 * V(Y) / (DYN[JP + 1] * DYP[JP]);
        r.b = -ddz_m / (DZP[KP - 1] * DZN[KP]);
        r.t = -ddz_p / (DZP[KP] * DZN[KP]);
    }
    else
    {
        const double ddx_p = 0.25 * (((V(0) + V(Z)) + V(X)) + V(X + Z));
        const double ddx_m = 0.25 * (((V(-X) + V(Z - X)) + V(0)) + V(Z));
        if(!two)
        {
            const double ddy_p = 0.25 * (((V(0) + V(Z)) + V(Y)) + V(Y + Z));
            const double ddy_m = 0.25 * (((V(-Y) + V(Z - Y)) + V(0)) + V(Z));
            r.p = 2.0 * V(Z) / (DZN[KP + 1] * DZP[KP]) + 2.0 * V(0) / (DZN[KP] * DZP[KP]) + ddx_p / (DXP[IP] * DXN[IP])
                + ddx_m / (DXP[IP - 1] * DXN[IP]) + ddy_p / (DYP[JP] * DYN[JP]) + ddy_m / (DYP[JP - 1] * DYN[JP])
                + CPOR / (alpha * dt);
            r.rhs = 0.0 + (((au[c + Z] - u[c]) * ddx_p - (u[c - X + Z] - u[c - X]) * ddx_m) / (DZP[KP] * DXN[IP])
                           + ((av[c + Z] - v[c]) * ddy_p - (v[c - Y + Z] - v[c - Y]) * ddy_m) / (DZP[KP] * DYN[JP])
                           + (CPOR * vin[c]) / (alpha * dt));
            r.e = -ddy_m / (DYP[JP - 1] * DYN[JP]);
            r.w = -ddy_p / (DYP[JP] * DYN[JP]);
        }
        else
        {
            r.p = 2.0 * V(Z) / (DZN[KP + 1] * DZP[KP]) + 2.0 * V(0) / (DZN[KP] * DZP[KP]) + ddx_p / (DXP[IP] * DXN[IP])
                + ddx_m / (DXP[IP - 1] * DXN[IP]) + CPOR / (alpha * dt);
            r.rhs = 0.0 + (((u[c + Z] - u[c]) * ddx_p - (u[c - X + Z] - u[c - X]) * ddx_m) / (DZP[KP] * DXN[IP])
                           + (CPOR * vin[c]) / (alpha * dt));
        }
        r.s = -ddx_m / (DXP[IP - 1] * DXN[IP]);
        r.n = -ddx_p / (DXP[IP] * DXN[IP]);
        r.b = -2.0 * V(0) / (DZN[KP] * DZP[KP]);
        r.t = -2.0 * V(Z) / (DZN[KP + 1] * DZP[KP]);
    }
    if(df)
    {
        const double x = (comp == 1 ? u : comp == 2 ? v : w)[c];
        if(i - 1 < 0) { r.rhs -= r.s * x; r.s = 0.0; }
        if(i + 1 >= ie) { r.rhs -= r.n * x; r.n = 0.0; }
        if(!two)
        {
            if(j - 1 < 0) { r.rhs -= r.e * x; r.e = 0.0; }
            if(j + 1 >= je) { r.rhs -= r.w * x; r.w = 0.0; }
        }
        if(k - 1 < 0) { r.rhs -= r.b * x; r.b = 0.0; }
        if(k + 1 >= ke) { r.rhs -= r.t * x; r.t = 0.0; }
    }
    return r;
}

}  // namespace reef
