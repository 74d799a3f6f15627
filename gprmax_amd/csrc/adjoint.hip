// Differentiable FDTD for per-cell permittivity and conductivity: a forward
// solve that keeps checkpoints, and the discrete adjoint (reverse mode) of it,
// which accumulates dL/dCA, dL/dCB and dL/dIEA per cell for a trace-misfit
// loss.  C ABI in include/gpx.h (gpx_inv_*).
//
// The forward scheme is the reference's solver loop (gprMax/model_build_run.py:
// 424-470) restricted to what a per-cell inversion needs: Yee updates
// (gprMax/fields_updates_ext.pyx:81-107, 404-412) with every cell its own
// material (E edge (c, i, j, k) takes cell (i, j, k), the reference's
// build_voxel without averaging), mu_r = 1 and no magnetic loss (DA = 1,
// DB = dt / (mu0 d)), HORIPML with one CFS term on any set of slabs
// (gprMax/pml_updates/pml_updates_*_HORIPML_ext.pyx, order1_*), optional
// Debye poles per cell with real auxiliary values (dispersive parts A and B,
// gprMax/fields_updates_ext.pyx:113-235; coefficients of
// gprMax/materials.py:92-143), one z-directed Hertzian dipole per solve
// (gprMax/sources.py:163-193), and Ez receivers sampled at the start of
// each iteration (gprMax/fields_outputs.py:40-67).
//
// With poles, an updated E edge of a cell with coefficients (ee, eqt, zt)
// per pole p does, around its E update:
//   A: phi = sum_p ee_p T_p;  T_p = eqt_p T_p + zt_p E^n;
//      E^{n+1} = CA E^n + CB curl H^{n+1} - IEA phi  (+ PML, source)
//   B: T_p = T_p - zt_p E^{n+1}
// so T^{n+1}_p = eqt_p T^n_p + zt_p (E^n - E^{n+1}).  Its reverse, with lT_p
// = dL/dT^{n+1}_p: the adjoint of E^{n+1} seen by the E update is
// lE* = lE - sum_p zt_p lT_p (k_adj_disp, before the gathers, which then
// use lE*); then lE^n = CA lE* + sum_p zt_p lT_p, lT^n_p = eqt_p lT_p -
// IEA ee_p lE*, and dL/dIEA gains -lE* phi^n (phi^n saved by the recompute).
//
// Iteration n maps the state s^n = (E^n, H^n, PhiE^n, PhiH^n) to s^{n+1}
// linearly.  Its adjoint, applied from n = N-1 down to 0 to the adjoint state
// (lE, lH, mE, mH) = dL/ds^{n+1}, is two gather kernels:
//   k_adj_e : lH += (dE^{n+1}/dH^{n+1})^T lE  (curl and E-PML terms); gradients
//             from lE and the stored E^n and curl H^{n+1}; mE = (dE/dPhiE)^T lE + RE mE;
//             lE = CA lE
//   k_adj_h : lE += (dH^{n+1}/dE^n)^T lH      (curl and H-PML terms);
//             mH = (dH/dPhiH)^T lH + RE mH; lH = DA lH = lH; and the
//             receiver residual of iteration n is added to lEz at each
//             receiver (the thread of the point adds it after its update).
// lE, mE and mH
// are double-buffered (read one copy, write the other), so each kernel's
// gathers see the values of the previous sub-step only.  Forward states
// come from checkpoints every K iterations; each segment is recomputed
// forward, storing E^n and curl H^{n+1} for every iteration of it, before the
// adjoint sweeps back over it.

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "gpx.h"

namespace {

constexpr int INV_MAX_SLABS = 6;
constexpr int INV_MAX_POLES = 4;

// One PML slab of the inversion grid (order 1).  Boxes are the corrected
// cells of each field; the profile index along the normal is p = pend - c
// (minus slabs) or c - pend (plus slabs).
struct InvSlab {
    int axis, minus;
    int loE[3], hiE[3], loH[3], hiH[3];
    int pendE, pendH;
    int org[3];
    int dE1[3], dE2[3], dH1[3], dH2[3];  // Phi dims (without the order axis)
    const float *RE;                     // E: RA, RB, RE, RF, each [t]
    const float *RH;                     // H: the same
    int t;
    float *PE1, *PE2, *PH1, *PH2;        // Phi (forward) or their adjoints (read)
    float *QE1, *QE2, *QH1, *QH2;        // adjoint: the adjoints written
    // Points whose adjoint gathers see this slab: the E box and the box one
    // step down the normal (E), the H box and the box one step up (H).
    int gE_lo[3], gE_hi[3], gH_lo[3], gH_hi[3];
};

struct InvArgs {
    int nx, ny, nz;
    long long ntot;
    float d, DB, cH;                     // cell size; DB = dt/(mu0 d); cH = dt/mu0
    float *E[3], *H[3];                  // fields (forward) or adjoint fields
    // Adjoint fields, one float4 per grid point: lE of the sub-step (read),
    // lE^n (written by the E half, completed by the H half) and lH.  The w of
    // lE and lE^n is CB of the point's cell, that of lH its CA (0 where the
    // point is no cell), carried along so that the E gather's CB lE terms and
    // CA come with the loads it makes anyway.
    float4 *lE, *lEn, *lH;
    float *EC;                           // forward: where to save, per point, E^n (3 floats)
                                         // then the curl of H^{n+1} (3), as three float2
                                         // (or null); adjoint: the saved values
    const float *CA, *CB, *IEA;          // per cell [nx][ny][nz]
    float2 *gCAB;                        // gradient accumulators (adjoint): dL/dCA, dL/dCB per cell
    float *gIEA;                         // and dL/dIEA
    int nslab;
    InvSlab slab[INV_MAX_SLABS];
    int si, sj, sk;                      // source edge (Ez), -1 when none
    float sfac;                          // dl / (dx dy dz)
    const float *waves;                  // [iterations]
    int it, iterations;
    long long ncell;
    int P;                               // Debye poles (0: none)
    const float *dc;                     // [3][P][ncell]: ee, eqt, zt of every cell and pole
    float *T[3];                         // forward: T of E component c, [P][ntot]; adjoint: lT
    float *Ps[3];                        // forward: where to save phi^n (or null); adjoint: phi^n
    int nrx;
    const int *rxc;                      // [nrx][3]
    const uint8_t *rxrow;                // adjoint: per (i, j) row of points, 1 when a receiver is on it
    float *traces;                       // [nrx][iterations]
    const float *resid;                  // [nrx][iterations], dL/d(trace)
    // Grid points [fast[0], fast[1]) x [fast[2], fast[3]) x [fast[4], fast[5])
    // have every neighbour of the adjoint gathers inside the grid and updated,
    // and no PML term: the adjoint kernels take a short path there.
    int fast[6];
    int shell[6][6];                     // [lo x, y, z, hi x, y, z] of the shell boxes
    int mix_nf, mix_ns, mix_gf[2], mix_gs;  // block counts and grid sizes of the mixed launches
    // Launches over every grid point: all_gk blocks of 128 points along k per
    // (j, i) row, all_blocks blocks, padded to a multiple of 8 (all_per_xcd x 8).
    int all_gk, all_blocks, all_per_xcd;
};

// One-dimensional launches: the hardware deals block b to XCD b mod 8, so block
// b takes logical number (b mod 8) * per_xcd + b / 8 (per_xcd = launched blocks
// / 8).  Each XCD then works through a contiguous run of logical blocks, i.e.
// of neighbouring grid rows, and the rows a point's differences read (j +- 1,
// i +- 1) are in the same XCD's L2 instead of another's.
__device__ __forceinline__ int xcd_block(int per_xcd)
{
    const int b = blockIdx.x;
    return (b & 7) * per_xcd + (b >> 3);
}

// Point (i, j, k) of a launch over every grid point; false past the grid.
__device__ __forceinline__ bool all_point(const InvArgs &a, int &i, int &j, int &k)
{
    const int lb = xcd_block(a.all_per_xcd);
    if (lb >= a.all_blocks) return false;
    const int kb = lb % a.all_gk, r = lb / a.all_gk;
    k = kb * blockDim.x + threadIdx.x;
    j = r % (a.ny + 1);
    i = r / (a.ny + 1);
    return true;
}

// Component c (0, 1, 2) of an adjoint field's float4.
__device__ __forceinline__ float comp3(const float4 &v, int c)
{
    return c == 0 ? v.x : (c == 1 ? v.y : v.z);
}

// Grid offsets are 32-bit: gpx_inv_create requires fewer than 2^31 points.
__device__ __forceinline__ int fidx(const InvArgs &a, int i, int j, int k)
{
    return (i * (a.ny + 1) + j) * (a.nz + 1) + k;
}
__device__ __forceinline__ int cidx(const InvArgs &a, int i, int j, int k)
{
    return (i * a.ny + j) * a.nz + k;
}
__device__ __forceinline__ bool in_box3(const int *lo, const int *hi, int i, int j, int k)
{
    return i >= lo[0] && i < hi[0] && j >= lo[1] && j < hi[1] && k >= lo[2] && k < hi[2];
}

// Reference update ranges (gprMax/fields_updates_ext.pyx): component c of E or H
// is updated at (i, j, k).
__device__ __forceinline__ bool e_upd(const InvArgs &a, int c, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0) return false;
    if (c == 0) return i < a.nx && j >= 1 && j < a.ny && k >= 1 && k < a.nz;
    if (c == 1) return i >= 1 && i < a.nx && j < a.ny && k >= 1 && k < a.nz;
    return i >= 1 && i < a.nx && j >= 1 && j < a.ny && k < a.nz;
}
__device__ __forceinline__ bool h_upd(const InvArgs &a, int c, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0) return false;
    if (c == 0) return i >= 1 && i <= a.nx && j < a.ny && k < a.nz;
    if (c == 1) return i < a.nx && j >= 1 && j <= a.ny && k < a.nz;
    return i < a.nx && j < a.ny && k >= 1 && k <= a.nz;
}

__device__ __forceinline__ bool in_e(const InvSlab &s, int i, int j, int k)
{
    return i >= s.loE[0] && i < s.hiE[0] && j >= s.loE[1] && j < s.hiE[1] && k >= s.loE[2] && k < s.hiE[2];
}
__device__ __forceinline__ bool in_h(const InvSlab &s, int i, int j, int k)
{
    return i >= s.loH[0] && i < s.hiH[0] && j >= s.loH[1] && j < s.hiH[1] && k >= s.loH[2] && k < s.hiH[2];
}

// Profile index and Phi index of cell (i, j, k) for the given field's box.
__device__ __forceinline__ void pidx(const InvSlab &s, bool efield, int i, int j, int k, int &p, int &q1,
                                     int &q2)
{
    int c[3] = {i - s.org[0], j - s.org[1], k - s.org[2]};
    const int along = s.axis == 0 ? i : (s.axis == 1 ? j : k);
    const int pend = efield ? s.pendE : s.pendH;
    p = s.minus ? pend - along : along - pend;
    c[s.axis] = p;
    const int *d1 = efield ? s.dE1 : s.dH1, *d2 = efield ? s.dE2 : s.dH2;
    q1 = (c[0] * d1[1] + c[1]) * d1[2] + c[2];
    q2 = (c[0] * d2[1] + c[1]) * d2[2] + c[2];
}

// E-PML pairs per axis: (corrected E component, source H component, sign) for
// Phi1 and Phi2; H-PML likewise (gprMax/pml_updates/*_HORIPML_ext.pyx).
__device__ __forceinline__ void e_pair(int axis, int w, int &comp, int &src, int &sign)
{
    const int tab[3][2][3] = {{{1, 2, -1}, {2, 1, +1}}, {{0, 2, +1}, {2, 0, -1}}, {{0, 1, -1}, {1, 0, +1}}};
    comp = tab[axis][w][0];
    src = tab[axis][w][1];
    sign = tab[axis][w][2];
}
__device__ __forceinline__ void h_pair(int axis, int w, int &comp, int &src, int &sign)
{
    const int tab[3][2][3] = {{{1, 2, +1}, {2, 1, -1}}, {{0, 2, -1}, {2, 0, +1}}, {{0, 1, +1}, {1, 0, -1}}};
    comp = tab[axis][w][0];
    src = tab[axis][w][1];
    sign = tab[axis][w][2];
}

__device__ __forceinline__ int axis_of(const InvSlab &s, int i, int j, int k)
{
    return s.axis == 0 ? i : (s.axis == 1 ? j : k);
}

// ------------------------------------------------------------------ forward
// H update and H-PML of one iteration; one thread per grid point.
__global__ void k_inv_h(InvArgs a)
{
    {
        int i, j, k;
        if (!all_point(a, i, j, k) || k > a.nz) return;
        const int g = fidx(a, i, j, k);
        const int o = g, oi = fidx(a, i + 1, j, k), oj = fidx(a, i, j + 1, k), ok = o + 1;
        float h[3] = {a.H[0][o], a.H[1][o], a.H[2][o]};
        if (h_upd(a, 0, i, j, k))
            h[0] = 1.0f * h[0] - a.DB * (a.E[2][oj] - a.E[2][o]) + a.DB * (a.E[1][ok] - a.E[1][o]);
        if (h_upd(a, 1, i, j, k))
            h[1] = 1.0f * h[1] - a.DB * (a.E[0][ok] - a.E[0][o]) + a.DB * (a.E[2][oi] - a.E[2][o]);
        if (h_upd(a, 2, i, j, k))
            h[2] = 1.0f * h[2] - a.DB * (a.E[1][oi] - a.E[1][o]) + a.DB * (a.E[0][oj] - a.E[0][o]);
        for (int s = 0; s < a.nslab; ++s) {
            const InvSlab &S = a.slab[s];
            if (!in_h(S, i, j, k)) continue;
            int p;
            int q[2];
            pidx(S, false, i, j, k, p, q[0], q[1]);
            const float RA01 = S.RH[p] - 1.0f, RB0 = S.RH[S.t + p], RE0 = S.RH[2 * S.t + p],
                        RF0 = S.RH[3 * S.t + p];
            const int on = S.axis == 0 ? oi : (S.axis == 1 ? oj : ok);
            const float ph2[2] = {S.PH1[q[0]], S.PH2[q[1]]};  // both loads before either store
            for (int w = 0; w < 2; ++w) {
                int comp, src, sign;
                h_pair(S.axis, w, comp, src, sign);
                float *Phi = w == 0 ? S.PH1 : S.PH2;
                const float dd = (a.E[src][on] - a.E[src][o]) / a.d;
                const float ph = ph2[w];
                const float corr = a.cH * (RA01 * dd + RB0 * ph);
                Phi[q[w]] = RE0 * ph - RF0 * dd;
                h[comp] = sign > 0 ? h[comp] + corr : h[comp] - corr;
            }
        }
        a.H[0][o] = h[0];
        a.H[1][o] = h[1];
        a.H[2][o] = h[2];
    }
}

// Receiver samples of iteration it (the state before its H update).
__global__ void k_inv_rx(InvArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.nrx) return;
    const int *c = a.rxc + 3 * r;
    a.traces[(long long)r * a.iterations + a.it] = a.E[2][fidx(a, c[0], c[1], c[2])];
}

// E update, E-PML and the Hertzian dipole of one iteration.
__global__ void k_inv_e(InvArgs a)
{
    {
        int i, j, k;
        if (!all_point(a, i, j, k) || k > a.nz) return;
        const int g = fidx(a, i, j, k);
        const int o = g;
        const int om[3] = {i > 0 ? fidx(a, i - 1, j, k) : o, j > 0 ? fidx(a, i, j - 1, k) : o, k > 0 ? o - 1 : o};
        float e[3] = {a.E[0][o], a.E[1][o], a.E[2][o]};
        const float en[3] = {e[0], e[1], e[2]};  // E^n, saved at the end (stores after every load)
        // The differences of H^{n+1} of the E update (om clamps to the point
        // itself at the low faces, so every load is inside the grid).
        const float h0 = a.H[0][o], h1 = a.H[1][o], h2 = a.H[2][o];
        const float d2y = h2 - a.H[2][om[1]], d1z = h1 - a.H[1][om[2]];
        const float d0z = h0 - a.H[0][om[2]], d2x = h2 - a.H[2][om[0]];
        const float d1x = h1 - a.H[1][om[0]], d0y = h0 - a.H[0][om[1]];
        const bool cell = i < a.nx && j < a.ny && k < a.nz;
        const int oc = cell ? cidx(a, i, j, k) : 0;
        const float CA = cell ? a.CA[oc] : 0.f, CB = cell ? a.CB[oc] : 0.f, IEA = cell ? a.IEA[oc] : 0.f;
        const bool upd[3] = {e_upd(a, 0, i, j, k), e_upd(a, 1, i, j, k), e_upd(a, 2, i, j, k)};
        // dispersive part A: phi from T^n, then T advanced with E^n
        float phi[3] = {0.f, 0.f, 0.f}, tv[3][INV_MAX_POLES];
        const bool disp = a.P > 0 && cell;
        if (disp) {
            for (int c = 0; c < 3; ++c) {
                if (!upd[c]) continue;
                float ph = 0.f;
                for (int p = 0; p < a.P; ++p) {
                    const float ee = a.dc[(long long)p * a.ncell + oc];
                    const float eq = a.dc[(long long)(a.P + p) * a.ncell + oc];
                    const float zt = a.dc[(long long)(2 * a.P + p) * a.ncell + oc];
                    const float t = a.T[c][(long long)p * a.ntot + o];
                    ph = ph + ee * t;
                    tv[c][p] = eq * t + zt * e[c];
                }
                phi[c] = ph;
            }
        }
        if (upd[0]) e[0] = CA * e[0] + CB * d2y - CB * d1z;
        if (upd[1]) e[1] = CA * e[1] + CB * d0z - CB * d2x;
        if (upd[2]) e[2] = CA * e[2] + CB * d1x - CB * d0y;
        if (disp)
            for (int c = 0; c < 3; ++c)
                if (upd[c]) e[c] = e[c] - IEA * phi[c];
        for (int s = 0; s < a.nslab; ++s) {
            const InvSlab &S = a.slab[s];
            if (!in_e(S, i, j, k)) continue;
            int p;
            int q[2];
            pidx(S, true, i, j, k, p, q[0], q[1]);
            const float RA01 = S.RE[p] - 1.0f, RB0 = S.RE[S.t + p], RE0 = S.RE[2 * S.t + p],
                        RF0 = S.RE[3 * S.t + p];
            const float ph2[2] = {S.PE1[q[0]], S.PE2[q[1]]};  // both loads before either store
            for (int w = 0; w < 2; ++w) {
                int comp, src, sign;
                e_pair(S.axis, w, comp, src, sign);
                float *Phi = w == 0 ? S.PE1 : S.PE2;
                const float dd = (a.H[src][o] - a.H[src][om[S.axis]]) / a.d;
                const float ph = ph2[w];
                const float corr = IEA * (RA01 * dd + RB0 * ph);
                Phi[q[w]] = RE0 * ph - RF0 * dd;
                e[comp] = sign > 0 ? e[comp] + corr : e[comp] - corr;
            }
        }
        if (i == a.si && j == a.sj && k == a.sk)
            e[2] = e[2] - IEA * a.waves[a.it] * a.sfac;
        if (disp)  // dispersive part B
            for (int c = 0; c < 3; ++c) {
                if (!upd[c]) continue;
                for (int p = 0; p < a.P; ++p) {
                    const float zt = a.dc[(long long)(2 * a.P + p) * a.ncell + oc];
                    a.T[c][(long long)p * a.ntot + o] = tv[c][p] - zt * e[c];
                }
            }
        a.E[0][o] = e[0];
        a.E[1][o] = e[1];
        a.E[2][o] = e[2];
        if (a.EC) {  // E^n, then the curl of H^{n+1} in the form dL/dCB reads it
            float2 *q = reinterpret_cast<float2 *>(a.EC) + 3 * (size_t)o;
            q[0] = make_float2(en[0], en[1]);
            q[1] = make_float2(en[2], d2y - d1z);
            q[2] = make_float2(d0z - d2x, d1x - d0y);
        }
        if (disp && a.Ps[0])
            for (int c = 0; c < 3; ++c) a.Ps[c][o] = phi[c];
    }
}

// ------------------------------------------------------------------ adjoint
// lE, lH, mE, mH live in a.E, a.H and the slab Phi pointers.

// dL/d(dd) of one E-PML term at E edge (comp of the pair w) of slab S; zero
// outside the slab's E box.
__device__ __forceinline__ float e_pml_ldd(const InvArgs &a, const InvSlab &S, int w, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0 || i > a.nx || j > a.ny || k > a.nz || !in_e(S, i, j, k)) return 0.f;
    int comp, src, sign;
    e_pair(S.axis, w, comp, src, sign);
    int p;
    int q[2];
    pidx(S, true, i, j, k, p, q[0], q[1]);
    const float RA01 = S.RE[p] - 1.0f, RF0 = S.RE[3 * S.t + p];
    const float IEA = (i < a.nx && j < a.ny && k < a.nz) ? a.IEA[cidx(a, i, j, k)] : 0.f;
    const float lF = comp3(a.lE[fidx(a, i, j, k)], comp);
    const float mu = (w == 0 ? S.PE1 : S.PE2)[q[w]];
    return (float)sign * IEA * RA01 * lF - RF0 * mu;
}

__device__ __forceinline__ float h_pml_ldd(const InvArgs &a, const InvSlab &S, int w, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0 || i > a.nx || j > a.ny || k > a.nz || !in_h(S, i, j, k)) return 0.f;
    int comp, src, sign;
    h_pair(S.axis, w, comp, src, sign);
    int p;
    int q[2];
    pidx(S, false, i, j, k, p, q[0], q[1]);
    const float RA01 = S.RH[p] - 1.0f, RF0 = S.RH[3 * S.t + p];
    const float lF = comp3(a.lH[fidx(a, i, j, k)], comp);
    const float mu = (w == 0 ? S.PH1 : S.PH2)[q[w]];
    return (float)sign * a.cH * RA01 * lF - RF0 * mu;
}

// Reverse of the dispersive parts of updated E edge (c, o) of cell oc, given
// lE* = dL/dE^{n+1} with the B part folded in (k_adj_disp): lE^n gains
// sum_p zt_p lT_p, lT_p becomes eqt_p lT_p - IEA ee_p lE*, and dL/dIEA
// gains -lE* phi^n.
__device__ __forceinline__ void disp_reverse(const InvArgs &a, int c, int o, int oc, float IEA, float lEs,
                                             float &out, float &giea)
{
    float add = 0.f;
    for (int p = 0; p < a.P; ++p) {
        const float ee = a.dc[(long long)p * a.ncell + oc];
        const float eq = a.dc[(long long)(a.P + p) * a.ncell + oc];
        const float zt = a.dc[(long long)(2 * a.P + p) * a.ncell + oc];
        float *lt = a.T[c] + (long long)p * a.ntot + o;
        const float l = *lt;
        add = add + zt * l;
        *lt = eq * l - IEA * ee * lEs;
    }
    out = out + add;
    giea += -lEs * a.Ps[c][o];
}

// lE* = lE - sum_p zt_p lT_p on the updated E edges of cells with poles: the
// adjoint of E^{n+1} as the E update sees it (B subtracts zt_p E^{n+1} from
// T^{n+1}).  Runs before the E gathers of the same iteration.
__global__ void k_adj_disp(InvArgs a)
{
    int i, j, k;
    if (!all_point(a, i, j, k) || k >= a.nz || j >= a.ny || i >= a.nx) return;
    const int o = fidx(a, i, j, k), oc = cidx(a, i, j, k);
    const float4 v = a.lE[o];
    float l3[3] = {v.x, v.y, v.z};
    for (int c = 0; c < 3; ++c) {
        if (!e_upd(a, c, i, j, k)) continue;
        float l = l3[c];
        for (int p = 0; p < a.P; ++p)
            l = l - a.dc[(long long)(2 * a.P + p) * a.ncell + oc] * a.T[c][(long long)p * a.ntot + o];
        l3[c] = l;
    }
    a.lE[o] = make_float4(l3[0], l3[1], l3[2], v.w);
}

// The adjoint fields at the start of the backward: zero, with CB of the
// point's cell in the w of both lE copies and its CA in the w of lH (0 where
// the point is no cell).
__global__ void k_adj_init(InvArgs a)
{
    int i, j, k;
    if (!all_point(a, i, j, k) || k > a.nz) return;
    const int o = fidx(a, i, j, k);
    const bool cell = i < a.nx && j < a.ny && k < a.nz;
    const float cb = cell ? a.CB[cidx(a, i, j, k)] : 0.f, ca = cell ? a.CA[cidx(a, i, j, k)] : 0.f;
    a.lE[o] = make_float4(0.f, 0.f, 0.f, cb);
    a.lEn[o] = make_float4(0.f, 0.f, 0.f, cb);
    a.lH[o] = make_float4(0.f, 0.f, 0.f, ca);
}

// lEz^n at point (i, j, k) after the H half: v plus dL/d(trace sample of
// iteration it) of every receiver on the point, in receiver order.
__device__ __forceinline__ float adj_rx_add(const InvArgs &a, int i, int j, int k, float v)
{
    if (!a.rxrow[i * (a.ny + 1) + j]) return v;
    for (int r = 0; r < a.nrx; ++r) {
        const int *c = a.rxc + 3 * r;
        if (c[0] == i && c[1] == j && c[2] == k) v += a.resid[(long long)r * a.iterations + a.it];
    }
    return v;
}

// CB of the cell of E edge (i, j, k) times lE_c there, when that edge is updated.
__device__ __forceinline__ float xi(const InvArgs &a, int c, int i, int j, int k)
{
    if (!e_upd(a, c, i, j, k)) return 0.f;
    const float4 v = a.lE[fidx(a, i, j, k)];  // w: CB of the cell
    return v.w * comp3(v, c);
}

// Reverse of the E half of iteration n, one thread per grid point:
//   lH    += (dE^{n+1}/dH^{n+1})^T lE^{n+1}      (curl and E-PML terms, a gather)
//   lE^n   = CA lE^{n+1} on updated edges (else unchanged), into En
//   mE^n   = RE mE^{n+1} + (dE^{n+1}/dPhiE^n)^T lE^{n+1}, into the Q arrays
//   gradients of the cell from lE^{n+1}, the saved E^n and curl H^{n+1}.
// lE^{n+1} and mE^{n+1} are only read, so neighbours see consistent values.
// FAST = true runs over the points of the fast box only (every lane of a wave
// takes the short path); FAST = false over all other points.
template <bool FAST>
__device__ __forceinline__ void adj_e_point(const InvArgs &a, int bx, int by)
{
    {
        // launched over (segments of a (j, k) plane, i) of the fast box (FAST),
        // every lane a point of the plane, or over the shell boxes
        int i, j, k;
        if (FAST) {
            const int fz = a.fast[5] - a.fast[4], t = bx * blockDim.x + threadIdx.x;
            if (t >= (a.fast[3] - a.fast[2]) * fz) return;
            k = a.fast[4] + t % fz;
            j = a.fast[2] + t / fz;
            i = a.fast[0] + by;
        } else {
            // the shell around the fast box, as up to six disjoint boxes
            // (blockIdx.y), each enumerated flat so every lane has a point
            const int *lo = a.shell[by], *hi = a.shell[by] + 3;
            const int sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
            const int t = bx * blockDim.x + threadIdx.x;
            if (sx <= 0 || sy <= 0 || sz <= 0 || t >= sx * sy * sz) return;
            k = lo[2] + t % sz;
            j = lo[1] + (t / sz) % sy;
            i = lo[0] + t / (sz * sy);
        }
        const int g = fidx(a, i, j, k);
        if (FAST) {
            // The same arithmetic with direct offsets: every edge involved is
            // updated and no PML term exists here.
            const int o = g;
            const int NZ1 = a.nz + 1, PL = (a.ny + 1) * NZ1;
            const int oc = cidx(a, i, j, k);
            // CB lE at the point and at its k+1, j+1 and i+1 neighbours (the
            // w of each is CB of that point's cell).
            // Every load comes before the first store: the compiler cannot
            // move a load above a store through these pointers, and a load
            // after a store costs the point a second memory round trip.
            const float4 q0 = a.lE[o], qk = a.lE[o + 1], qj = a.lE[o + NZ1], qi = a.lE[o + PL];
            const float dh0 = q0.w * q0.y - qk.w * qk.y - q0.w * q0.z + qj.w * qj.z;
            const float dh1 = q0.w * q0.z - qi.w * qi.z - q0.w * q0.x + qk.w * qk.x;
            const float dh2 = q0.w * q0.x - qj.w * qj.x - q0.w * q0.y + qi.w * qi.y;
            const float lE[3] = {q0.x, q0.y, q0.z};
            const float4 hq = a.lH[o];  // w: CA of the cell
            const float CA = hq.w;
            const float2 *sq = reinterpret_cast<const float2 *>(a.EC) + 3 * (size_t)o;
            const float2 s0 = sq[0], s1 = sq[1], s2 = sq[2];
            const float En[3] = {s0.x, s0.y, s1.x}, curl[3] = {s1.y, s2.x, s2.y};
            float gca = 0.f, gcb = 0.f;
            for (int c = 0; c < 3; ++c) {
                gca += lE[c] * En[c];
                gcb += lE[c] * curl[c];
            }
            const float2 gab = a.gCAB[oc];
            const float gca0 = gab.x, gcb0 = gab.y;
            a.lH[o] = make_float4(hq.x + dh0, hq.y + dh1, hq.z + dh2, hq.w);
            a.gCAB[oc] = make_float2(gca0 + gca, gcb0 + gcb);
            float giea = 0.f;
            if (i == a.si && j == a.sj && k == a.sk) giea += -lE[2] * a.waves[a.it] * a.sfac;
            float out[3] = {CA * lE[0], CA * lE[1], CA * lE[2]};
            if (a.P > 0) {
                const float IEA = a.IEA[oc];
                for (int c = 0; c < 3; ++c) disp_reverse(a, c, o, oc, IEA, lE[c], out[c], giea);
            }
            if (giea != 0.f) a.gIEA[oc] += giea;
            a.lEn[o] = make_float4(out[0], out[1], out[2], q0.w);
            return;
        }
        // Ex(i,j,k) = ... + CB (Hz(i,j,k) - Hz(i,j-1,k)) - CB (Hy(i,j,k) - Hy(i,j,k-1))
        // Ey(i,j,k) = ... + CB (Hx(i,j,k) - Hx(i,j,k-1)) - CB (Hz(i,j,k) - Hz(i-1,j,k))
        // Ez(i,j,k) = ... + CB (Hy(i,j,k) - Hy(i-1,j,k)) - CB (Hx(i,j,k) - Hx(i,j-1,k))
        float dh[3];
        if (i >= 1 && i + 2 <= a.nx && j >= 1 && j + 2 <= a.ny && k >= 1 && k + 2 <= a.nz) {
            // every edge the curl terms read is updated: xi without its range
            // checks, as on the fast path (same terms, same order)
            const int o = g, NZ1 = a.nz + 1, PL = (a.ny + 1) * NZ1;
            const float4 q0 = a.lE[o], qk = a.lE[o + 1], qj = a.lE[o + NZ1], qi = a.lE[o + PL];
            dh[0] = q0.w * q0.y - qk.w * qk.y - q0.w * q0.z + qj.w * qj.z;
            dh[1] = q0.w * q0.z - qi.w * qi.z - q0.w * q0.x + qk.w * qk.x;
            dh[2] = q0.w * q0.x - qj.w * qj.x - q0.w * q0.y + qi.w * qi.y;
        } else {
            dh[0] = xi(a, 1, i, j, k) - xi(a, 1, i, j, k + 1) - xi(a, 2, i, j, k) + xi(a, 2, i, j + 1, k);
            dh[1] = xi(a, 2, i, j, k) - xi(a, 2, i + 1, j, k) - xi(a, 0, i, j, k) + xi(a, 0, i, j, k + 1);
            dh[2] = xi(a, 0, i, j, k) - xi(a, 0, i, j + 1, k) - xi(a, 1, i, j, k) + xi(a, 1, i + 1, j, k);
        }
        const int o = g;
        // Loads first (see the fast path), then the stores.
        const float4 le = a.lE[o];
        float lE[3] = {le.x, le.y, le.z};
        float out[3] = {lE[0], lE[1], lE[2]};
        const float4 hq = a.lH[o];  // w: CA of the cell (0 where the point is no cell)
        const bool cell = i < a.nx && j < a.ny && k < a.nz;
        float gca = 0.f, gcb = 0.f, gca0 = 0.f, gcb0 = 0.f;
        int oc = 0;
        float CA = 0.f, IEA = 0.f;
        if (cell) {
            oc = cidx(a, i, j, k);
            const int om[3] = {i > 0 ? fidx(a, i - 1, j, k) : o, j > 0 ? fidx(a, i, j - 1, k) : o,
                                     k > 0 ? o - 1 : o};
            CA = hq.w;
            IEA = a.IEA[oc];
            const float2 *sq = reinterpret_cast<const float2 *>(a.EC) + 3 * (size_t)o;
            const float2 s0 = sq[0], s1 = sq[1], s2 = sq[2];
            const float En[3] = {s0.x, s0.y, s1.x}, curl[3] = {s1.y, s2.x, s2.y};
            for (int c = 0; c < 3; ++c) {
                if (!e_upd(a, c, i, j, k)) continue;
                gca += lE[c] * En[c];
                gcb += lE[c] * curl[c];
            }
            const float2 gab = a.gCAB[oc];
            gca0 = gab.x;
            gcb0 = gab.y;
        }
        // E-PML: dd = (S(e) - S(e - axis)) / d uses H at the edge and at edge - axis,
        // so H here feeds the edge here (+) and the edge one step up the axis (-).
        // The edge's own term (e_pml_ldd of the edge, with its lE and IEA, 0
        // outside the cells) shares the slab index, profile values and Phi
        // adjoints with the update of the Phi adjoints, mE^n = RE mE^{n+1} +
        // (dE/dPhiE)^T lE, written into the Q arrays (edges of cells only;
        // edges outside the cells are never updated).
        for (int s = 0; s < a.nslab; ++s) {
            const InvSlab &S = a.slab[s];
            if (!in_box3(S.gE_lo, S.gE_hi, i, j, k)) continue;
            const bool own = in_e(S, i, j, k);
            int p = 0;
            int q[2] = {0, 0};
            float mu[2] = {0.f, 0.f}, RA01 = 0.f, RB0 = 0.f, RE0 = 0.f, RF0 = 0.f;
            if (own) {
                pidx(S, true, i, j, k, p, q[0], q[1]);
                RA01 = S.RE[p] - 1.0f;
                RB0 = S.RE[S.t + p];
                RE0 = S.RE[2 * S.t + p];
                RF0 = S.RE[3 * S.t + p];
                mu[0] = S.PE1[q[0]];
                mu[1] = S.PE2[q[1]];
            }
            const int ip = i + (S.axis == 0), jp = j + (S.axis == 1), kp = k + (S.axis == 2);
            for (int w = 0; w < 2; ++w) {
                int comp, src, sign;
                e_pair(S.axis, w, comp, src, sign);
                const float ldd = own ? (float)sign * IEA * RA01 * lE[comp] - RF0 * mu[w] : 0.f;
                dh[src] += (ldd - e_pml_ldd(a, S, w, ip, jp, kp)) / a.d;
                if (own && cell) (w == 0 ? S.QE1 : S.QE2)[q[w]] = RE0 * mu[w] + (float)sign * IEA * RB0 * lE[comp];
            }
        }
        a.lH[o] = make_float4(hq.x + dh[0], hq.y + dh[1], hq.z + dh[2], hq.w);
        if (cell) {  // edges outside the cells are never updated
            float giea = 0.f;
            for (int c = 0; c < 3; ++c) {
                if (!e_upd(a, c, i, j, k)) continue;
                out[c] = CA * lE[c];
                if (a.P > 0) disp_reverse(a, c, o, oc, IEA, lE[c], out[c], giea);
            }
            if (i == a.si && j == a.sj && k == a.sk) giea += -lE[2] * a.waves[a.it] * a.sfac;
            a.gCAB[oc] = make_float2(gca0 + gca, gcb0 + gcb);
            if (giea != 0.f) a.gIEA[oc] += giea;
        }
        a.lEn[o] = make_float4(out[0], out[1], out[2], le.w);
    }
}

template <bool FAST>
__global__ void k_adj_e(InvArgs a)
{
    adj_e_point<FAST>(a, blockIdx.x, blockIdx.y);
}

// DB times lH_c at H cell (i, j, k), when that cell is updated.
__device__ __forceinline__ float eta(const InvArgs &a, int c, int i, int j, int k)
{
    if (!h_upd(a, c, i, j, k)) return 0.f;
    return a.DB * comp3(a.lH[fidx(a, i, j, k)], c);
}

// Reverse of the H half of iteration n (lH = dL/dH^{n+1}, complete):
//   lE^n  += (dH^{n+1}/dE^n)^T lH                  (curl and H-PML terms, a gather)
//   mH^n   = RE mH^{n+1} + (dH^{n+1}/dPhiH^n)^T lH, into the Q arrays
// and lH^n = DA lH = lH.
// FAST = true runs over the points of the fast box only (every lane of a wave
// takes the short path); FAST = false over all other points.
template <bool FAST>
__device__ __forceinline__ void adj_h_point(const InvArgs &a, int bx, int by)
{
    {
        // launched over (segments of a (j, k) plane, i) of the fast box (FAST),
        // every lane a point of the plane, or over the shell boxes
        int i, j, k;
        if (FAST) {
            const int fz = a.fast[5] - a.fast[4], t = bx * blockDim.x + threadIdx.x;
            if (t >= (a.fast[3] - a.fast[2]) * fz) return;
            k = a.fast[4] + t % fz;
            j = a.fast[2] + t / fz;
            i = a.fast[0] + by;
        } else {
            // the shell around the fast box, as up to six disjoint boxes
            // (blockIdx.y), each enumerated flat so every lane has a point
            const int *lo = a.shell[by], *hi = a.shell[by] + 3;
            const int sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
            const int t = bx * blockDim.x + threadIdx.x;
            if (sx <= 0 || sy <= 0 || sz <= 0 || t >= sx * sy * sz) return;
            k = lo[2] + t % sz;
            j = lo[1] + (t / sz) % sy;
            i = lo[0] + t / (sz * sy);
        }
        const int g = fidx(a, i, j, k);
        if (FAST) {
            const int o = g;
            const int NZ1 = a.nz + 1, PL = (a.ny + 1) * NZ1;
            // DB lH at the point and at its k-1, j-1 and i-1 neighbours
            const float4 h0 = a.lH[o], hk = a.lH[o - 1], hj = a.lH[o - NZ1], hi = a.lH[o - PL];
            const float DB = a.DB;
            const float de0 = DB * h0.y - DB * hk.y - DB * h0.z + DB * hj.z;
            const float de1 = DB * h0.z - DB * hi.z - DB * h0.x + DB * hk.x;
            const float de2 = DB * h0.x - DB * hj.x - DB * h0.y + DB * hi.y;
            const float4 e = a.lE[o];
            a.lE[o] = make_float4(e.x + de0, e.y + de1, adj_rx_add(a, i, j, k, e.z + de2), e.w);
            return;
        }
        // Hx(i,j,k) = Hx - DB (Ez(i,j+1,k) - Ez(i,j,k)) + DB (Ey(i,j,k+1) - Ey(i,j,k))
        // Hy(i,j,k) = Hy - DB (Ex(i,j,k+1) - Ex(i,j,k)) + DB (Ez(i+1,j,k) - Ez(i,j,k))
        // Hz(i,j,k) = Hz - DB (Ey(i+1,j,k) - Ey(i,j,k)) + DB (Ex(i,j+1,k) - Ex(i,j,k))
        float de[3];
        if (i >= 1 && i + 1 <= a.nx && j >= 1 && j + 1 <= a.ny && k >= 1 && k + 1 <= a.nz) {
            // every cell the curl terms read is updated: eta without its range
            // checks, as on the fast path (same terms, same order)
            const int o = g, NZ1 = a.nz + 1, PL = (a.ny + 1) * NZ1;
            const float4 h0 = a.lH[o], hk = a.lH[o - 1], hj = a.lH[o - NZ1], hi = a.lH[o - PL];
            const float DB = a.DB;
            de[0] = DB * h0.y - DB * hk.y - DB * h0.z + DB * hj.z;
            de[1] = DB * h0.z - DB * hi.z - DB * h0.x + DB * hk.x;
            de[2] = DB * h0.x - DB * hj.x - DB * h0.y + DB * hi.y;
        } else {
            de[0] = eta(a, 1, i, j, k) - eta(a, 1, i, j, k - 1) - eta(a, 2, i, j, k) + eta(a, 2, i, j - 1, k);
            de[1] = eta(a, 2, i, j, k) - eta(a, 2, i - 1, j, k) - eta(a, 0, i, j, k) + eta(a, 0, i, j, k - 1);
            de[2] = eta(a, 0, i, j, k) - eta(a, 0, i, j - 1, k) - eta(a, 1, i, j, k) + eta(a, 1, i - 1, j, k);
        }
        // Loads first (lH here and lE for the update), then the stores.
        const float4 lh = a.lH[g];
        const float lH[3] = {lh.x, lh.y, lh.z};
        const float4 ev = a.lE[g];
        // H-PML: dd = (S(h + axis) - S(h)) / d, so E here feeds the cell here (-)
        // and the cell one step down the axis (+).  The cell's own term
        // (h_pml_ldd of the cell, with its lH) shares the slab index, profile
        // values and Phi adjoints with the update of the Phi adjoints, mH^n =
        // RE mH^{n+1} + (dH/dPhiH)^T lH, written into the Q arrays.
        for (int s = 0; s < a.nslab; ++s) {
            const InvSlab &S = a.slab[s];
            if (!in_box3(S.gH_lo, S.gH_hi, i, j, k)) continue;
            const bool own = in_h(S, i, j, k);
            int p = 0;
            int q[2] = {0, 0};
            float mu[2] = {0.f, 0.f}, RA01 = 0.f, RB0 = 0.f, RE0 = 0.f, RF0 = 0.f;
            if (own) {
                pidx(S, false, i, j, k, p, q[0], q[1]);
                RA01 = S.RH[p] - 1.0f;
                RB0 = S.RH[S.t + p];
                RE0 = S.RH[2 * S.t + p];
                RF0 = S.RH[3 * S.t + p];
                mu[0] = S.PH1[q[0]];
                mu[1] = S.PH2[q[1]];
            }
            const int im = i - (S.axis == 0), jm = j - (S.axis == 1), km = k - (S.axis == 2);
            for (int w = 0; w < 2; ++w) {
                int comp, src, sign;
                h_pair(S.axis, w, comp, src, sign);
                const float ldd = own ? (float)sign * a.cH * RA01 * lH[comp] - RF0 * mu[w] : 0.f;
                de[src] += (h_pml_ldd(a, S, w, im, jm, km) - ldd) / a.d;
                if (own) (w == 0 ? S.QH1 : S.QH2)[q[w]] = RE0 * mu[w] + (float)sign * a.cH * RB0 * lH[comp];
            }
        }
        a.lE[g] = make_float4(ev.x + de[0], ev.y + de[1], adj_rx_add(a, i, j, k, ev.z + de[2]), ev.w);
    }
}

template <bool FAST>
__global__ void k_adj_h(InvArgs a)
{
    adj_h_point<FAST>(a, blockIdx.x, blockIdx.y);
}

// The fast-box and shell blocks of one half-step in one launch: the shell
// blocks (latency-bound gathers) are spread evenly among the fast ones
// (bandwidth-bound) so that the two kinds run side by side.  Block b is a
// shell block when floor((b+1) ns / (nf+ns)) > floor(b ns / (nf+ns)).
__device__ __forceinline__ bool mix_shell(const InvArgs &a, int b, int &idx)
{
    const long long tot = (long long)a.mix_nf + a.mix_ns;
    const long long c0 = (long long)b * a.mix_ns / tot, c1 = (long long)(b + 1) * a.mix_ns / tot;
    idx = c1 > c0 ? (int)c0 : (int)(b - c1);
    return c1 > c0;
}

__global__ void k_adj_e_mix(InvArgs a)
{
    int q;
    if (mix_shell(a, blockIdx.x, q)) adj_e_point<false>(a, q % a.mix_gs, q / a.mix_gs);
    else adj_e_point<true>(a, q % a.mix_gf[0], q / a.mix_gf[0]);
}

__global__ void k_adj_h_mix(InvArgs a)
{
    int q;
    if (mix_shell(a, blockIdx.x, q)) adj_h_point<false>(a, q % a.mix_gs, q / a.mix_gs);
    else adj_h_point<true>(a, q % a.mix_gf[0], q / a.mix_gf[0]);
}


}  // namespace

// ====================================================================== host
struct gpx_inv {
    gpx_inv_desc desc;
    long long ntot = 0, ncell = 0;
    hipStream_t stream = nullptr;
    float *F = nullptr;            // E, H: 6 x ntot
    float4 *LE[2] = {nullptr, nullptr};  // adjoint lE, two copies, and lH: one float4
    float4 *LH = nullptr;                // per grid point each (InvArgs::lE, lH)
    float *ck = nullptr;           // checkpoints: nck x (6 ntot + phi)
    float *seg = nullptr;          // segment store: K x 6 x ntot
    float *par = nullptr;          // CA, CB, IEA: 3 x ncell
    float *grad = nullptr;         // (gCA, gCB) interleaved per cell, then gIEA: 3 x ncell
    float *waves = nullptr;
    int P = 0;                     // Debye poles
    float *dc = nullptr;           // ee, eqt, zt: 3 x P x ncell
    float *T = nullptr;            // T of Ex, Ey, Ez: 3 x P x ntot
    float *LT = nullptr;           // its adjoint
    int *rxc = nullptr;
    uint8_t *rxrow = nullptr;      // [nx+1][ny+1]: 1 on the rows of points that hold a receiver
    float *traces = nullptr, *resid = nullptr;
    int nrx = 0, si = -1, sj = -1, sk = -1;
    float sfac = 0;
    bool have_model = false, have_forward = false;
    struct Slab {
        InvSlab s;
        long long n[4] = {0, 0, 0, 0};   // Phi sizes
        float *phi[4] = {nullptr, nullptr, nullptr, nullptr};
        float *mu[4] = {nullptr, nullptr, nullptr, nullptr};
        float *mu2[4] = {nullptr, nullptr, nullptr, nullptr};  // second copies
        float *R = nullptr;              // E then H tables, 8 t floats
    };
    std::vector<Slab> slabs;
    long long phi_total = 0;
    int fast[6] = {0, 0, 0, 0, 0, 0};  // see InvArgs
    int shell[6][6] = {};              // see InvArgs
    int shell_max = 0;                 // points of the largest shell box
    float ms_fwd = 0, ms_bwd = 0;
    // Segments whose stored states the forward keeps (gpx_inv_set_keep): the
    // last nkeep segments, each K x seg_stride floats, in `keep`; the backward
    // then skips their recompute.
    double keep_bytes = 0;
    float *keep = nullptr;
    long long keep_floats = 0;     // size of `keep`
    int nkeep = 0;                 // segments kept by the last forward
    std::string err;
};

static std::string g_inv_err;

static int ifail(gpx_inv *v, int code, const std::string &m)
{
    if (v) v->err = m;
    else g_inv_err = m;
    return code;
}

#define INV_TRY(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return ifail(v, (int)e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *gpx_inv_last_error(const gpx_inv *v)
{
    return v ? v->err.c_str() : g_inv_err.c_str();
}

// The box of grid points where the adjoint kernels need no range check and
// no PML term: points 1..n-2 on every axis, shrunk past every slab's E box
// and the edges one below it, and its H box and the cells one above it.
static void update_fast(gpx_inv *v)
{
    const int n[3] = {v->desc.nx, v->desc.ny, v->desc.nz};
    int lo[3] = {1, 1, 1}, hi[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
    for (const auto &S : v->slabs) {
        const InvSlab &s = S.s;
        const int ax = s.axis;
        if (s.minus) {
            const int end = s.hiE[ax] > s.hiH[ax] + 1 ? s.hiE[ax] : s.hiH[ax] + 1;
            if (end > lo[ax]) lo[ax] = end;
        } else {
            const int start = s.loE[ax] - 1 < s.loH[ax] ? s.loE[ax] - 1 : s.loH[ax];
            if (start < hi[ax]) hi[ax] = start;
        }
    }
    for (int q = 0; q < 3; ++q) {
        v->fast[2 * q] = lo[q];
        v->fast[2 * q + 1] = hi[q] > lo[q] ? hi[q] : lo[q];
    }
    // The grid points [0, n] minus the fast box: below / above it in x (all
    // y, z), then in y within its x range (all z), then in z within its x
    // and y ranges.
    const int N[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
    const int *f = v->fast;
    const int boxes[6][6] = {
        {0, 0, 0, f[0], N[1], N[2]},          {f[1], 0, 0, N[0], N[1], N[2]},
        {f[0], 0, 0, f[1], f[2], N[2]},       {f[0], f[3], 0, f[1], N[1], N[2]},
        {f[0], f[2], 0, f[1], f[3], f[4]},    {f[0], f[2], f[5], f[1], f[3], N[2]}};
    memcpy(v->shell, boxes, sizeof(boxes));
    v->shell_max = 0;
    for (auto &b : boxes) {
        const long long vol = (long long)(b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]);
        if (vol > v->shell_max) v->shell_max = (int)vol;
    }
}

extern "C" int gpx_inv_create(const gpx_inv_desc *d, gpx_inv **out)
{
    if (!d || !out) return ifail(nullptr, GPX_EARG, "gpx_inv_create: null argument");
    *out = nullptr;
    if (d->nx < 2 || d->ny < 2 || d->nz < 2 || d->iterations < 1 || d->K < 1 || !(d->d > 0) || !(d->dt > 0))
        return ifail(nullptr, GPX_EARG, "gpx_inv_create: bad descriptor");
    if (d->nx + 1 > 65535 || d->ny + 1 > 65535)
        return ifail(nullptr, GPX_ELIMIT, "gpx_inv_create: at most 65534 cells along x and y");
    if ((double)(d->nx + 1) * (d->ny + 1) * (d->nz + 1) >= 2147483647.0)
        return ifail(nullptr, GPX_ELIMIT, "gpx_inv_create: at most 2^31 grid points");
    if (hipSetDevice(d->device) != hipSuccess) return ifail(nullptr, GPX_EARG, "gpx_inv_create: bad device");
    gpx_inv *v = new gpx_inv();
    v->desc = *d;
    v->ntot = (long long)(d->nx + 1) * (d->ny + 1) * (d->nz + 1);
    v->ncell = (long long)d->nx * d->ny * d->nz;
    hipError_t e = hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&v->F, 6 * v->ntot * 4);
    for (float4 **p : {&v->LE[0], &v->LE[1], &v->LH})
        if (e == hipSuccess) e = hipMalloc((void **)p, v->ntot * 16);
    if (e == hipSuccess) e = hipMalloc((void **)&v->seg, (size_t)d->K * ((6 * v->ntot + 3) / 4 * 4) * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&v->par, 3 * v->ncell * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&v->grad, 3 * v->ncell * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&v->waves, (size_t)d->iterations * 4);
    if (e != hipSuccess) {
        gpx_inv_destroy(v);
        return ifail(nullptr, (int)e, std::string("gpx_inv_create: ") + hipGetErrorString(e));
    }
    update_fast(v);
    *out = v;
    return 0;
}

extern "C" int gpx_inv_destroy(gpx_inv *v)
{
    if (!v) return 0;
    hipSetDevice(v->desc.device);
    if (v->stream) hipStreamSynchronize(v->stream);
    for (float4 *p : {v->LE[0], v->LE[1], v->LH})
        if (p) hipFree(p);
    for (float *p : {v->F, v->ck, v->seg, v->par, v->grad, v->waves, v->traces, v->resid, v->dc,
                     v->T, v->LT, v->keep})
        if (p) hipFree(p);
    if (v->rxc) hipFree(v->rxc);
    if (v->rxrow) hipFree(v->rxrow);
    for (auto &S : v->slabs) {
        for (float *p : S.phi) if (p) hipFree(p);
        for (float *p : S.mu) if (p) hipFree(p);
        for (float *p : S.mu2) if (p) hipFree(p);
        if (S.R) hipFree(S.R);
    }
    if (v->stream) hipStreamDestroy(v->stream);
    delete v;
    return 0;
}

extern "C" int gpx_inv_add_pml(gpx_inv *v, int dir, int xs, int xf, int ys, int yf, int zs, int zf,
                               const float *eR, const float *hR)
{
    if (!v || !eR || !hR || dir < 0 || dir > 5) return ifail(v, GPX_EARG, "gpx_inv_add_pml: bad argument");
    if ((int)v->slabs.size() >= INV_MAX_SLABS) return ifail(v, GPX_ELIMIT, "gpx_inv_add_pml: at most 6 slabs");
    const gpx_inv_desc &g = v->desc;
    if (xs < 0 || ys < 0 || zs < 0 || xf > g.nx || yf > g.ny || zf > g.nz || xs >= xf || ys >= yf || zs >= zf)
        return ifail(v, GPX_EARG, "gpx_inv_add_pml: slab box outside the grid");
    hipSetDevice(g.device);
    gpx_inv::Slab S;
    InvSlab &s = S.s;
    memset(&s, 0, sizeof(s));
    const int axis = dir % 3, minus = dir < 3;
    const int st[3] = {xs, ys, zs}, en[3] = {xf, yf, zf};
    const int t = en[axis] - st[axis];
    s.axis = axis;
    s.minus = minus;
    s.t = t;
    for (int n = 0; n < 3; ++n) {
        s.loE[n] = s.loH[n] = st[n];
        s.hiE[n] = s.hiH[n] = en[n];
        s.org[n] = st[n];
    }
    if (false) {
    } else if (minus) {  // E cells f - p, H cells f - (p + 1)
        s.loE[axis] = en[axis] - t + 1;
        s.hiE[axis] = en[axis] + 1;
        s.pendE = en[axis];
        s.loH[axis] = en[axis] - t;
        s.hiH[axis] = en[axis];
        s.pendH = en[axis] - 1;
    } else {
        s.loE[axis] = s.loH[axis] = st[axis];
        s.hiE[axis] = s.hiH[axis] = st[axis] + t;
        s.pendE = s.pendH = st[axis];
    }
    for (int q = 0; q < 3; ++q) {
        s.gE_lo[q] = s.loE[q] - (q == axis);
        s.gE_hi[q] = s.hiE[q];
        s.gH_lo[q] = s.loH[q];
        s.gH_hi[q] = s.hiH[q] + (q == axis);
    }
    // Phi shapes of gprMax/pml.py:202-219 (order 1)
    const int nx = xf - xs, ny = yf - ys, nz = zf - zs;
    int dims[4][3];
    if (axis == 0) {
        const int d4[4][3] = {{nx + 1, ny, nz + 1}, {nx + 1, ny + 1, nz}, {nx, ny + 1, nz}, {nx, ny, nz + 1}};
        memcpy(dims, d4, sizeof(dims));
    } else if (axis == 1) {
        const int d4[4][3] = {{nx, ny + 1, nz + 1}, {nx + 1, ny + 1, nz}, {nx + 1, ny, nz}, {nx, ny, nz + 1}};
        memcpy(dims, d4, sizeof(dims));
    } else {
        const int d4[4][3] = {{nx, ny + 1, nz + 1}, {nx + 1, ny, nz + 1}, {nx + 1, ny, nz}, {nx, ny + 1, nz}};
        memcpy(dims, d4, sizeof(dims));
    }
    for (int n = 0; n < 3; ++n) {
        s.dE1[n] = dims[0][n];
        s.dE2[n] = dims[1][n];
        s.dH1[n] = dims[2][n];
        s.dH2[n] = dims[3][n];
    }
    for (int q = 0; q < 4; ++q) {
        S.n[q] = (long long)dims[q][0] * dims[q][1] * dims[q][2];
        INV_TRY(hipMalloc((void **)&S.phi[q], S.n[q] * 4));
        INV_TRY(hipMalloc((void **)&S.mu[q], S.n[q] * 4));
        INV_TRY(hipMalloc((void **)&S.mu2[q], S.n[q] * 4));
        v->phi_total += S.n[q];
    }
    INV_TRY(hipMalloc((void **)&S.R, (size_t)8 * t * 4));
    INV_TRY(hipMemcpy(S.R, eR, (size_t)4 * t * 4, hipMemcpyHostToDevice));
    INV_TRY(hipMemcpy(S.R + 4 * t, hR, (size_t)4 * t * 4, hipMemcpyHostToDevice));
    s.RE = S.R;
    s.RH = S.R + 4 * t;
    v->slabs.push_back(S);
    update_fast(v);
    if (v->ck) {  // checkpoint layout depends on the Phi sizes
        hipFree(v->ck);
        v->ck = nullptr;
    }
    v->have_forward = false;
    return 0;
}

extern "C" int gpx_inv_set_model(gpx_inv *v, const float *CA, const float *CB, const float *IEA)
{
    if (!v || !CA || !CB || !IEA) return ifail(v, GPX_EARG, "gpx_inv_set_model: null argument");
    hipSetDevice(v->desc.device);
    INV_TRY(hipMemcpy(v->par, CA, v->ncell * 4, hipMemcpyHostToDevice));
    INV_TRY(hipMemcpy(v->par + v->ncell, CB, v->ncell * 4, hipMemcpyHostToDevice));
    INV_TRY(hipMemcpy(v->par + 2 * v->ncell, IEA, v->ncell * 4, hipMemcpyHostToDevice));
    v->have_model = true;
    v->have_forward = false;
    return 0;
}

extern "C" int gpx_inv_set_dispersive(gpx_inv *v, int P, const float *ee, const float *eqt, const float *zt)
{
    if (!v || P < 0 || P > INV_MAX_POLES || (P > 0 && (!ee || !eqt || !zt)))
        return ifail(v, GPX_EARG, "gpx_inv_set_dispersive: bad argument (at most 4 poles)");
    hipSetDevice(v->desc.device);
    if (P != v->P) {
        for (float **p : {&v->dc, &v->T, &v->LT, &v->ck, &v->seg}) {
            if (*p) hipFree(*p);
            *p = nullptr;
        }
        v->P = P;
        // the segment store keeps phi^n of the three E components beside E^n, curl H^{n+1}
        INV_TRY(hipMalloc((void **)&v->seg, (size_t)v->desc.K * (((P ? 9 : 6) * v->ntot + 3) / 4 * 4) * 4));
        if (P) {
            INV_TRY(hipMalloc((void **)&v->dc, (size_t)3 * P * v->ncell * 4));
            INV_TRY(hipMalloc((void **)&v->T, (size_t)3 * P * v->ntot * 4));
            INV_TRY(hipMalloc((void **)&v->LT, (size_t)3 * P * v->ntot * 4));
        }
    }
    const size_t b = (size_t)P * v->ncell * 4;
    if (P) {
        INV_TRY(hipMemcpy(v->dc, ee, b, hipMemcpyHostToDevice));
        INV_TRY(hipMemcpy(v->dc + (size_t)P * v->ncell, eqt, b, hipMemcpyHostToDevice));
        INV_TRY(hipMemcpy(v->dc + (size_t)2 * P * v->ncell, zt, b, hipMemcpyHostToDevice));
    }
    v->have_forward = false;
    return 0;
}

static int n_checkpoints(const gpx_inv *v) { return (v->desc.iterations + v->desc.K - 1) / v->desc.K; }
static long long ck_stride(const gpx_inv *v) { return 6 * v->ntot + 3 * v->P * v->ntot + v->phi_total; }
// Floats per iteration in the segment store: E^n and curl H^{n+1} interleaved
// (6 per point), then phi^n of the three E components with poles; rounded up
// to 4 floats so that every slot is 16-byte aligned.
static long long seg_stride(const gpx_inv *v) { return ((v->P ? 9 : 6) * v->ntot + 3) / 4 * 4; }

static int all_gk(const gpx_inv *v);
static int all_blocks(const gpx_inv *v);

static InvArgs inv_args(const gpx_inv *v, bool adjoint, int it)
{
    InvArgs a;
    memset(&a, 0, sizeof(a));
    const gpx_inv_desc &g = v->desc;
    a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
    a.ntot = v->ntot;
    a.d = g.d;
    const double mu0 = 1.25663706127e-6;
    a.DB = (float)((1.0 / g.d) * (g.dt / mu0));
    a.cH = (float)(g.dt / mu0);
    for (int c = 0; c < 3 && !adjoint; ++c) {  // the adjoint fields are set by the backward
        a.E[c] = v->F + c * v->ntot;
        a.H[c] = v->F + (3 + c) * v->ntot;
    }
    a.CA = v->par;
    a.CB = v->par + v->ncell;
    a.IEA = v->par + 2 * v->ncell;
    a.gCAB = reinterpret_cast<float2 *>(v->grad);
    a.gIEA = v->grad + 2 * v->ncell;
    a.nslab = (int)v->slabs.size();
    for (int s = 0; s < a.nslab; ++s) {
        a.slab[s] = v->slabs[s].s;
        float *const *P = adjoint ? v->slabs[s].mu : v->slabs[s].phi;
        a.slab[s].PE1 = P[0];
        a.slab[s].PE2 = P[1];
        a.slab[s].PH1 = P[2];
        a.slab[s].PH2 = P[3];
    }
    a.si = v->si; a.sj = v->sj; a.sk = v->sk;
    a.sfac = v->sfac;
    a.waves = v->waves;
    a.it = it;
    a.iterations = g.iterations;
    a.nrx = v->nrx;
    a.rxc = v->rxc;
    a.rxrow = v->rxrow;
    a.traces = v->traces;
    a.resid = v->resid;
    a.ncell = v->ncell;
    a.P = v->P;
    a.dc = v->dc;
    for (int c = 0; c < 3 && v->P; ++c) a.T[c] = (adjoint ? v->LT : v->T) + (size_t)c * v->P * v->ntot;
    memcpy(a.fast, v->fast, sizeof(a.fast));
    memcpy(a.shell, v->shell, sizeof(a.shell));
    a.all_gk = all_gk(v);
    a.all_blocks = all_blocks(v);
    a.all_per_xcd = (a.all_blocks + 7) / 8;
    return a;
}

// Launch shapes: 128 threads along k, one block row per (j, i).  Launches over
// every point are one-dimensional (all_point): all_gk blocks per row, the
// block count padded to a multiple of 8.
static int all_gk(const gpx_inv *v)
{
    return (v->desc.nz + 1 + 127) / 128;
}
static int all_blocks(const gpx_inv *v)
{
    return all_gk(v) * (v->desc.ny + 1) * (v->desc.nx + 1);
}
static dim3 grid_all(const gpx_inv *v)
{
    return dim3((unsigned)((all_blocks(v) + 7) / 8 * 8));
}
// Threads per block of the adjoint gathers (fast-box and shell blocks).
constexpr int ADJ_BLOCK = 256;  // 64: 3.06x, 128: 2.99x, 256: 2.91x, 512: 2.98x, 1024: 3.12x (kept segments, 160x160x120)
static dim3 grid_fast(const gpx_inv *v)
{
    const int fx = v->fast[1] - v->fast[0], fy = v->fast[3] - v->fast[2], fz = v->fast[5] - v->fast[4];
    const long long plane = fy > 0 && fz > 0 ? (long long)fy * fz : 1;
    return dim3((unsigned)((plane + ADJ_BLOCK - 1) / ADJ_BLOCK), fx > 0 ? fx : 1, 1);
}
static dim3 grid_shell(const gpx_inv *v)
{
    return dim3(v->shell_max > 0 ? (v->shell_max + ADJ_BLOCK - 1) / ADJ_BLOCK : 1, 6, 1);
}
static bool fast_empty(const gpx_inv *v)
{
    return v->fast[1] <= v->fast[0] || v->fast[3] <= v->fast[2] || v->fast[5] <= v->fast[4];
}

// Copies the state (fields and Phi) to or from checkpoint c.
static int checkpoint(gpx_inv *v, int c, bool save)
{
    float *dst = v->ck + (long long)c * ck_stride(v);
    auto cp = [&](float *ckp, float *live, long long n) {
        return save ? hipMemcpyAsync(ckp, live, n * 4, hipMemcpyDeviceToDevice, v->stream)
                    : hipMemcpyAsync(live, ckp, n * 4, hipMemcpyDeviceToDevice, v->stream);
    };
    INV_TRY(cp(dst, v->F, 6 * v->ntot));
    long long off = 6 * v->ntot;
    if (v->P) {
        INV_TRY(cp(dst + off, v->T, 3 * v->P * v->ntot));
        off += 3 * v->P * v->ntot;
    }
    for (auto &S : v->slabs)
        for (int q = 0; q < 4; ++q) {
            INV_TRY(cp(dst + off, S.phi[q], S.n[q]));
            off += S.n[q];
        }
    return 0;
}

// One forward iteration; with `slot`, the kernels also store E^n, curl H^{n+1}
// (and phi^n) there, as the recompute does.
static void forward_step(gpx_inv *v, int it, bool rx, float *slot = nullptr)
{
    InvArgs a = inv_args(v, false, it);
    if (slot) a.EC = slot;
    for (int q = 0; slot && v->P && q < 3; ++q) a.Ps[q] = slot + (6 + q) * v->ntot;
    if (rx && v->nrx) hipLaunchKernelGGL(k_inv_rx, dim3((v->nrx + 63) / 64), dim3(64), 0, v->stream, a);
    hipLaunchKernelGGL(k_inv_h, grid_all(v), dim3(128), 0, v->stream, a);
    hipLaunchKernelGGL(k_inv_e, grid_all(v), dim3(128), 0, v->stream, a);
}

extern "C" int gpx_inv_forward(gpx_inv *v, int si, int sj, int sk, float sfac, const float *waves, int nrx,
                               const int *rx, float *traces)
{
    if (!v || !waves || nrx < 0 || (nrx > 0 && (!rx || !traces)))
        return ifail(v, GPX_EARG, "gpx_inv_forward: bad argument");
    if (!v->have_model) return ifail(v, GPX_ESTATE, "gpx_inv_forward: no model uploaded");
    const gpx_inv_desc &g = v->desc;
    if (si < 1 || si >= g.nx || sj < 1 || sj >= g.ny || sk < 0 || sk >= g.nz)
        return ifail(v, GPX_EARG, "gpx_inv_forward: source edge outside the updated Ez range");
    for (int r = 0; r < nrx; ++r)
        if (rx[3 * r] < 0 || rx[3 * r] > g.nx || rx[3 * r + 1] < 0 || rx[3 * r + 1] > g.ny || rx[3 * r + 2] < 0 ||
            rx[3 * r + 2] > g.nz)
            return ifail(v, GPX_EARG, "gpx_inv_forward: receiver outside the grid");
    hipSetDevice(g.device);
    v->si = si; v->sj = sj; v->sk = sk;
    v->sfac = sfac;
    INV_TRY(hipMemcpy(v->waves, waves, (size_t)g.iterations * 4, hipMemcpyHostToDevice));
    if (nrx != v->nrx || !v->rxc) {
        if (v->rxc) hipFree(v->rxc);
        if (v->traces) hipFree(v->traces);
        if (v->resid) hipFree(v->resid);
        v->rxc = nullptr;
        v->traces = v->resid = nullptr;
        v->nrx = nrx;
        INV_TRY(hipMalloc((void **)&v->rxc, (size_t)(nrx > 0 ? nrx : 1) * 3 * 4));
        INV_TRY(hipMalloc((void **)&v->traces, (size_t)(nrx > 0 ? nrx : 1) * g.iterations * 4));
        INV_TRY(hipMalloc((void **)&v->resid, (size_t)(nrx > 0 ? nrx : 1) * g.iterations * 4));
    }
    if (nrx) INV_TRY(hipMemcpy(v->rxc, rx, (size_t)nrx * 3 * 4, hipMemcpyHostToDevice));
    {  // the rows of points that hold a receiver, for the adjoint H half
        const size_t rows = (size_t)(g.nx + 1) * (g.ny + 1);
        if (!v->rxrow) INV_TRY(hipMalloc((void **)&v->rxrow, rows));
        std::vector<uint8_t> flag(rows, 0);
        for (int r = 0; r < nrx; ++r) flag[(size_t)rx[3 * r] * (g.ny + 1) + rx[3 * r + 1]] = 1;
        INV_TRY(hipMemcpy(v->rxrow, flag.data(), rows, hipMemcpyHostToDevice));
    }
    if (!v->ck) INV_TRY(hipMalloc((void **)&v->ck, (size_t)n_checkpoints(v) * ck_stride(v) * 4));
    // the segments whose states this forward keeps
    const int nseg = n_checkpoints(v);
    const long long seg_floats = (long long)g.K * seg_stride(v);
    const double fit = v->keep_bytes / ((double)seg_floats * 4.0);
    const int nkeep = fit >= nseg ? nseg : (int)fit;
    if ((long long)nkeep * seg_floats != v->keep_floats) {
        if (v->keep) hipFree(v->keep);
        v->keep = nullptr;
        v->keep_floats = 0;
        v->nkeep = 0;
        v->have_forward = false;
        if (nkeep > 0) {
            const hipError_t e = hipMalloc((void **)&v->keep, (size_t)nkeep * seg_floats * 4);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                v->keep = nullptr;
                return ifail(v, (int)e, "gpx_inv_forward: cannot allocate " +
                                              std::to_string((double)nkeep * seg_floats * 4 / 1e9) +
                                              " GB for the kept segments (gpx_inv_set_keep)");
            }
            v->keep_floats = (long long)nkeep * seg_floats;
        }
    }
    v->nkeep = nkeep;
    INV_TRY(hipMemsetAsync(v->F, 0, 6 * v->ntot * 4, v->stream));
    if (v->P) INV_TRY(hipMemsetAsync(v->T, 0, (size_t)3 * v->P * v->ntot * 4, v->stream));
    for (auto &S : v->slabs)
        for (int q = 0; q < 4; ++q) INV_TRY(hipMemsetAsync(S.phi[q], 0, S.n[q] * 4, v->stream));
    hipEvent_t e0, e1;
    INV_TRY(hipEventCreate(&e0));
    INV_TRY(hipEventCreate(&e1));
    INV_TRY(hipEventRecord(e0, v->stream));
    for (int it = 0; it < g.iterations; ++it) {
        const int c = it / g.K;
        if (c >= nseg - nkeep) {  // a kept segment: its states, no checkpoint
            float *slot = v->keep + (long long)(c - (nseg - nkeep)) * seg_floats +
                          (long long)(it - c * g.K) * seg_stride(v);
            forward_step(v, it, true, slot);
            continue;
        }
        if (it % g.K == 0)
            if (int rc = checkpoint(v, c, true)) return rc;
        forward_step(v, it, true);
    }
    INV_TRY(hipGetLastError());
    INV_TRY(hipEventRecord(e1, v->stream));
    INV_TRY(hipEventSynchronize(e1));
    INV_TRY(hipEventElapsedTime(&v->ms_fwd, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (nrx) INV_TRY(hipMemcpy(traces, v->traces, (size_t)nrx * g.iterations * 4, hipMemcpyDeviceToHost));
    v->have_forward = true;
    return 0;
}

extern "C" int gpx_inv_backward(gpx_inv *v, const float *resid, float *gCA, float *gCB, float *gIEA)
{
    if (!v || !gCA || !gCB || !gIEA || (v->nrx > 0 && !resid)) return ifail(v, GPX_EARG, "gpx_inv_backward: bad argument");
    if (!v->have_forward) return ifail(v, GPX_ESTATE, "gpx_inv_backward: run gpx_inv_forward first");
    const gpx_inv_desc &g = v->desc;
    hipSetDevice(g.device);
    if (v->nrx) INV_TRY(hipMemcpy(v->resid, resid, (size_t)v->nrx * g.iterations * 4, hipMemcpyHostToDevice));
    {  // zero adjoint fields, CB and CA of the cells in the w components
        InvArgs a = inv_args(v, true, 0);
        a.lE = v->LE[0];
        a.lEn = v->LE[1];
        a.lH = v->LH;
        hipLaunchKernelGGL(k_adj_init, grid_all(v), dim3(128), 0, v->stream, a);
        INV_TRY(hipGetLastError());
    }
    INV_TRY(hipMemsetAsync(v->grad, 0, 3 * v->ncell * 4, v->stream));
    if (v->P) INV_TRY(hipMemsetAsync(v->LT, 0, (size_t)3 * v->P * v->ntot * 4, v->stream));
    for (auto &S : v->slabs)
        for (int q = 0; q < 4; ++q) {
            INV_TRY(hipMemsetAsync(S.mu[q], 0, S.n[q] * 4, v->stream));
            INV_TRY(hipMemsetAsync(S.mu2[q], 0, S.n[q] * 4, v->stream));
        }
    float4 *lEb[2] = {v->LE[0], v->LE[1]};  // current and next copy of lE
    int cur = 0;
    hipEvent_t e0, e1;
    INV_TRY(hipEventCreate(&e0));
    INV_TRY(hipEventCreate(&e1));
    INV_TRY(hipEventRecord(e0, v->stream));
    const dim3 G = grid_all(v), GF = grid_fast(v), GS = grid_shell(v);
    const bool nofast = fast_empty(v);
    const int nseg = n_checkpoints(v);
    for (int c = nseg - 1; c >= 0; --c) {
        const int n0 = c * g.K, n1 = n0 + g.K < g.iterations ? n0 + g.K : g.iterations;
        const bool kept = c >= nseg - v->nkeep;
        float *base = kept ? v->keep + (long long)(c - (nseg - v->nkeep)) * g.K * seg_stride(v) : v->seg;
        // recompute the segment, storing E^n and curl H^{n+1} (kept segments have them)
        if (!kept)
            if (int rc = checkpoint(v, c, false)) return rc;
        for (int it = n0; it < n1 && !kept; ++it) {
            float *slot = base + (long long)(it - n0) * seg_stride(v);
            InvArgs a = inv_args(v, false, it);
            a.EC = slot;  // the kernels save E^n, curl H^{n+1} (and phi^n) here
            for (int q = 0; v->P && q < 3; ++q) a.Ps[q] = slot + (6 + q) * v->ntot;
            hipLaunchKernelGGL(k_inv_h, G, dim3(128), 0, v->stream, a);
            hipLaunchKernelGGL(k_inv_e, G, dim3(128), 0, v->stream, a);
        }
        // adjoint sweep over the segment
        for (int it = n1 - 1; it >= n0; --it) {
            float *slot = base + (long long)(it - n0) * seg_stride(v);
            InvArgs a = inv_args(v, true, it);
            a.EC = slot;
            for (int q = 0; v->P && q < 3; ++q) a.Ps[q] = slot + (6 + q) * v->ntot;
            a.lE = lEb[cur];
            a.lEn = lEb[cur ^ 1];
            a.lH = v->LH;
            if (v->P) hipLaunchKernelGGL(k_adj_disp, G, dim3(128), 0, v->stream, a);
            for (int s = 0; s < a.nslab; ++s) {
                const auto &S = v->slabs[s];
                float *const *P = cur == 0 ? S.mu : S.mu2;
                float *const *Q = cur == 0 ? S.mu2 : S.mu;
                a.slab[s].PE1 = P[0]; a.slab[s].PE2 = P[1]; a.slab[s].PH1 = P[2]; a.slab[s].PH2 = P[3];
                a.slab[s].QE1 = Q[0]; a.slab[s].QE2 = Q[1]; a.slab[s].QH1 = Q[2]; a.slab[s].QH2 = Q[3];
            }
            if (nofast) {
                hipLaunchKernelGGL(k_adj_e<false>, GS, dim3(ADJ_BLOCK), 0, v->stream, a);
                a.lE = a.lEn;  // lE^n, completed by the H half
                hipLaunchKernelGGL(k_adj_h<false>, GS, dim3(ADJ_BLOCK), 0, v->stream, a);
            } else {
                a.mix_nf = (int)(GF.x * GF.y * GF.z);
                a.mix_ns = (int)(GS.x * GS.y);
                a.mix_gf[0] = (int)GF.x;
                a.mix_gf[1] = (int)GF.y;
                a.mix_gs = (int)GS.x;
                const dim3 GM(a.mix_nf + a.mix_ns);
                hipLaunchKernelGGL(k_adj_e_mix, GM, dim3(ADJ_BLOCK), 0, v->stream, a);
                a.lE = a.lEn;  // lE^n, completed by the H half
                hipLaunchKernelGGL(k_adj_h_mix, GM, dim3(ADJ_BLOCK), 0, v->stream, a);
            }
            cur ^= 1;
        }
        INV_TRY(hipGetLastError());
    }
    INV_TRY(hipEventRecord(e1, v->stream));
    INV_TRY(hipEventSynchronize(e1));
    INV_TRY(hipEventElapsedTime(&v->ms_bwd, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    std::vector<float> ab((size_t)2 * v->ncell);
    INV_TRY(hipMemcpy(ab.data(), v->grad, (size_t)2 * v->ncell * 4, hipMemcpyDeviceToHost));
    for (long long c = 0; c < v->ncell; ++c) {
        gCA[c] = ab[2 * c];
        gCB[c] = ab[2 * c + 1];
    }
    INV_TRY(hipMemcpy(gIEA, v->grad + 2 * v->ncell, v->ncell * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int gpx_inv_set_keep(gpx_inv *v, double bytes)
{
    if (!v || !(bytes >= 0)) return ifail(v, GPX_EARG, "gpx_inv_set_keep: bad argument");
    v->keep_bytes = bytes;
    return 0;
}

extern "C" int gpx_inv_timings(const gpx_inv *v, float *ms_forward, float *ms_backward)
{
    if (!v) return GPX_EARG;
    if (ms_forward) *ms_forward = v->ms_fwd;
    if (ms_backward) *ms_backward = v->ms_bwd;
    return 0;
}
