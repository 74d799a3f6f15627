// MI355X (gfx950) FDTD solve path: Yee E/H updates with fused CFS-PML slab
// corrections, dispersive multipole A/B, point sources, receivers and
// snapshots.  C ABI declared in include/gpx.h.
//
// Work mapping.  Field arrays are [nx+1][ny+1][nz+1], z fastest, so one x-plane
// is a contiguous run of (ny+1)(nz+1) floats.  A block owns `block` consecutive
// points of that plane and marches along x over a chunk of planes; each thread
// owns one (j, k) column.  Every load of a plane is therefore one coalesced
// stream, the x-neighbour is carried in registers from the previous plane (a
// register queue), and the y/z neighbours are the same stream shifted by
// (nz+1) or 1 elements and are served by L1/L2.  Blocks are remapped so that
// consecutive plane segments land on the same XCD and share its L2.
//
// Fusion.  The H pass applies, per cell and in the reference order: receiver
// store (old values), main H update, each H-PML slab in slab order, magnetic
// dipoles.  The E pass applies: main E update or dispersive part A, each E-PML
// slab in slab order, voltage sources, Hertzian dipoles, dispersive part B.
// Every correction of a cell is made by the thread that owns it, so no atomics
// are needed and the float32 rounding sequence is the CPU path's.
//
// Rounding.  Built with -ffp-contract=off and IEEE division, every expression
// rounds as written, in the reference's association order, so results are
// comparable bit for bit with a CPU run built the same way.

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "gpx.h"

namespace {

constexpr int MAX_SLABS = 6;
constexpr int MAX_POLES = 4;
constexpr int MAX_LDS_BYTES = 64 * 1024;

enum : int {
    F_MAIN = 1,
    F_SRC = 2,
    F_RX = 4,
    F_DISPB = 8,
    F_PML0 = 16,  // bits 4..9: one per slab
};
constexpr int F_PML_ALL = F_PML0 * ((1 << MAX_SLABS) - 1);

struct SlabArgs {
    int axis;        // 0 x, 1 y, 2 z
    int minus;       // 1 for xminus/yminus/zminus
    int lo[3], hi[3];// box of corrected cells for this field, [lo, hi)
    int pend;        // profile index p = pend - c (minus) or c - pend (plus)
    int org[3];      // slab origin (xs, ys, zs) for the Phi index
    int d1[3], d2[3];// Phi1/Phi2 dims without the order axis
    int t, order;
    float d;
    float *phi1, *phi2;
    const float *R;  // [4][order][t]
};

struct SrcArgs {
    int n;
    const int *info;     // [n][6]
    const float *fac;    // [n][2]
    const float *waves;  // [n][iterations]
};

struct StepArgs {
    int nx, ny, nz;
    int plane;           // (ny+1)*(nz+1)
    int nseg, nblocks, per_xcd, xc;
    float *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
    const uint32_t *ID;  // [6][ntot]
    long long ntot;
    const float *coef;   // [nmat][5]
    const float *disp;   // [nmat][3P] (real) or [nmat][3P][2] (complex)
    const uint8_t *dispflag;
    float *T[3];         // [P][ntot] or [P][ntot][2]
    int nmat;
    int nslab;
    SlabArgs slab[MAX_SLABS];
    int flags;
    int it, iterations;
    SrcArgs src0, src1;  // H: src0 magnetic; E: src0 voltage, src1 Hertzian
    int nrx;
    const int *rxc;      // [nrx][3]
    float *rxbuf;        // [6][iterations][nrx]
};

// ---------------------------------------------------------------- PML
// One corrected component (gprMax/pml_updates/*_ext.pyx).  R = [RA, RB, RE, RF]
// each [order][t]; idx/off index Phi order 0 and the order-1 offset.
template <int FORM>
__device__ __forceinline__ float pml_correct(const SlabArgs &s, int p, float F, bool plus_sign,
                                             float c, float dd, float *phi, long long idx,
                                             long long off)
{
    const int ot = s.order * s.t;
    const float *R = s.R;
    float corr;
    if (FORM == 0) {  // HORIPML
        if (s.order == 1) {
            float RA01 = R[p] - 1.0f;
            float RB0 = R[ot + p], RE0 = R[2 * ot + p], RF0 = R[3 * ot + p];
            float ph = phi[idx];
            corr = c * (RA01 * dd + RB0 * ph);
            phi[idx] = RE0 * ph - RF0 * dd;
        } else {
            float RA0 = R[p], RA1 = R[s.t + p];
            float RB0 = R[ot + p], RB1 = R[ot + s.t + p];
            float RE0 = R[2 * ot + p], RE1 = R[2 * ot + s.t + p];
            float RF0 = R[3 * ot + p], RF1 = R[3 * ot + s.t + p];
            float RA01 = RA0 * RA1 - 1.0f;
            float p0 = phi[idx], p1 = phi[idx + off];
            corr = c * (RA01 * dd + RA1 * RB0 * p0 + RB1 * p1);
            phi[idx + off] = RE1 * p1 - RF1 * (RA0 * dd + RB0 * p0);
            phi[idx] = RE0 * p0 - RF0 * dd;
        }
    } else {  // MRIPML
        if (s.order == 1) {
            float IRA = 1.0f / R[p];
            float IRA1 = IRA - 1.0f;
            float RB0 = R[ot + p], RE0 = R[2 * ot + p], RF0 = R[3 * ot + p];
            float RC0 = IRA * RB0 * RF0;
            float ph = phi[idx];
            corr = c * (IRA1 * dd - IRA * ph);
            phi[idx] = RE0 * ph + RC0 * dd - RC0 * ph;
        } else {
            float IRA = 1.0f / (R[p] + R[s.t + p]);
            float IRA1 = IRA - 1.0f;
            float RB0 = R[ot + p], RB1 = R[ot + s.t + p];
            float RE0 = R[2 * ot + p], RE1 = R[2 * ot + s.t + p];
            float RC0 = IRA * R[3 * ot + p], RC1 = IRA * R[3 * ot + s.t + p];
            float p0 = phi[idx], p1 = phi[idx + off];
            float psi = RB0 * p0 + RB1 * p1;
            corr = c * (IRA1 * dd - IRA * psi);
            phi[idx + off] = RE1 * p1 + RC1 * (dd - psi);
            phi[idx] = RE0 * p0 + RC0 * (dd - psi);
        }
    }
    return plus_sign ? F + corr : F - corr;
}

__device__ __forceinline__ bool in_box(const SlabArgs &s, int i, int j, int k)
{
    return i >= s.lo[0] && i < s.hi[0] && j >= s.lo[1] && j < s.hi[1] && k >= s.lo[2] &&
           k < s.hi[2];
}

// Profile index and the two Phi element indices of cell (i, j, k).
__device__ __forceinline__ void slab_index(const SlabArgs &s, int i, int j, int k, int &p,
                                           long long &i1, long long &i2, long long &o1,
                                           long long &o2)
{
    int c[3] = {i - s.org[0], j - s.org[1], k - s.org[2]};
    int along = s.axis == 0 ? i : (s.axis == 1 ? j : k);
    p = s.minus ? s.pend - along : along - s.pend;
    c[s.axis] = p;
    i1 = ((long long)c[0] * s.d1[1] + c[1]) * s.d1[2] + c[2];
    i2 = ((long long)c[0] * s.d2[1] + c[1]) * s.d2[2] + c[2];
    o1 = (long long)s.d1[0] * s.d1[1] * s.d1[2];
    o2 = (long long)s.d2[0] * s.d2[1] * s.d2[2];
}

// Block -> (plane segment, x chunk) with consecutive segments on one XCD.
__device__ __forceinline__ int logical_block(const StepArgs &a)
{
    int b = blockIdx.x;
    return (b & 7) * a.per_xcd + (b >> 3);
}

// ---------------------------------------------------------------- H pass
template <int FORM>
__global__ void k_h_pass(StepArgs a)
{
    extern __shared__ float lds[];
    const int l = logical_block(a);
    if (l >= a.nblocks) return;
    const int nc = a.nmat * 5;
    for (int n = threadIdx.x; n < nc; n += blockDim.x) lds[n] = a.coef[n];
    __syncthreads();

    const int seg = l % a.nseg, chunk = l / a.nseg;
    const int t = seg * blockDim.x + threadIdx.x;
    if (t >= a.plane) return;
    const int NZ1 = a.nz + 1;
    const int j = t / NZ1, k = t - j * NZ1;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = chunk * a.xc;
    const int i1 = min(i0 + a.xc, nx + 1);
    const long long P = a.plane, N = a.ntot;
    const int flags = a.flags;
    const uint32_t *ID = a.ID;

    long long f = (long long)i0 * P + t;
    float ey = a.Ey[f], ez = a.Ez[f];
    for (int i = i0; i < i1; ++i, f += P) {
        const float ex = a.Ex[f];
        float ey1 = 0.f, ez1 = 0.f;
        if (i < nx) {
            ey1 = a.Ey[f + P];
            ez1 = a.Ez[f + P];
        }
        float hx = a.Hx[f], hy = a.Hy[f], hz = a.Hz[f];

        if (flags & F_RX) {
            for (int r = 0; r < a.nrx; ++r) {
                if (a.rxc[3 * r] == i && a.rxc[3 * r + 1] == j && a.rxc[3 * r + 2] == k) {
                    float *o = a.rxbuf + (long long)a.it * a.nrx + r;
                    const long long cs = (long long)a.iterations * a.nrx;
                    o[0] = ex; o[cs] = ey; o[2 * cs] = ez;
                    o[3 * cs] = hx; o[4 * cs] = hy; o[5 * cs] = hz;
                }
            }
        }

        // Differences of E used by the H update and the H-PML (forward).
        const float dEz_y = (j < ny) ? a.Ez[f + NZ1] - ez : 0.f;
        const float dEy_z = (k < nz) ? a.Ey[f + 1] - ey : 0.f;
        const float dEx_z = (k < nz) ? a.Ex[f + 1] - ex : 0.f;
        const float dEz_x = (i < nx) ? ez1 - ez : 0.f;
        const float dEy_x = (i < nx) ? ey1 - ey : 0.f;
        const float dEx_y = (j < ny) ? a.Ex[f + NZ1] - ex : 0.f;
        const uint32_t m3 = ID[3 * N + f], m4 = ID[4 * N + f], m5 = ID[5 * N + f];

        if (flags & F_MAIN) {  // gprMax/fields_updates_ext.pyx:404-412
            if (i >= 1 && j < ny && k < nz)
                hx = lds[m3 * 5] * hx - lds[m3 * 5 + 2] * dEz_y + lds[m3 * 5 + 3] * dEy_z;
            if (i < nx && j >= 1 && k < nz)
                hy = lds[m4 * 5] * hy - lds[m4 * 5 + 3] * dEx_z + lds[m4 * 5 + 1] * dEz_x;
            if (i < nx && j < ny && k >= 1)
                hz = lds[m5 * 5] * hz - lds[m5 * 5 + 1] * dEy_x + lds[m5 * 5 + 2] * dEx_y;
        }

        if (flags & F_PML_ALL) {
            for (int s = 0; s < a.nslab; ++s) {
                if (!(flags & (F_PML0 << s))) continue;
                const SlabArgs &sl = a.slab[s];
                if (!in_box(sl, i, j, k)) continue;
                int p;
                long long q1, q2, o1, o2;
                slab_index(sl, i, j, k, p, q1, q2, o1, o2);
                if (sl.axis == 0) {
                    hy = pml_correct<FORM>(sl, p, hy, true, lds[m4 * 5 + 4], dEz_x / sl.d, sl.phi1, q1, o1);
                    hz = pml_correct<FORM>(sl, p, hz, false, lds[m5 * 5 + 4], dEy_x / sl.d, sl.phi2, q2, o2);
                } else if (sl.axis == 1) {
                    hx = pml_correct<FORM>(sl, p, hx, false, lds[m3 * 5 + 4], dEz_y / sl.d, sl.phi1, q1, o1);
                    hz = pml_correct<FORM>(sl, p, hz, true, lds[m5 * 5 + 4], dEx_y / sl.d, sl.phi2, q2, o2);
                } else {
                    hx = pml_correct<FORM>(sl, p, hx, true, lds[m3 * 5 + 4], dEy_z / sl.d, sl.phi1, q1, o1);
                    hy = pml_correct<FORM>(sl, p, hy, false, lds[m4 * 5 + 4], dEx_z / sl.d, sl.phi2, q2, o2);
                }
            }
        }

        if (flags & F_SRC) {  // magnetic dipoles, gprMax/sources.py:202-232
            for (int s = 0; s < a.src0.n; ++s) {
                const int *in = a.src0.info + 6 * s;
                if (in[0] == i && in[1] == j && in[2] == k && a.it >= in[4] && a.it <= in[5]) {
                    const int pol = in[3];
                    const uint32_t m = pol == 0 ? m3 : (pol == 1 ? m4 : m5);
                    const float w = a.src0.waves[(long long)s * a.iterations + a.it];
                    const float v = lds[m * 5 + 4] * w * a.src0.fac[2 * s];
                    if (pol == 0) hx = hx - v;
                    else if (pol == 1) hy = hy - v;
                    else hz = hz - v;
                }
            }
        }

        if (flags & (F_MAIN | F_PML_ALL | F_SRC)) {
            a.Hx[f] = hx;
            a.Hy[f] = hy;
            a.Hz[f] = hz;
        }
        ey = ey1;
        ez = ez1;
    }
}

// ---------------------------------------------------------------- E pass
// Dispersive part A for one component (gprMax/fields_updates_ext.pyx:113-296):
// returns phi and advances the P auxiliary values from the old field e.
template <int P, bool CPLX>
__device__ __forceinline__ float disp_a(const float *D, uint32_t m, float *tr, float *ti, float e)
{
    float phi = 0.f;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (CPLX) {
            const float *c = D + ((long long)m * 3 * P + 3 * p) * 2;
            if (P == 1) phi = c[0] * tr[p];
            else phi = phi + c[0] * tr[p];
            const float er = c[2], ei = c[3], zr = c[4], zi = c[5];
            const float nr = (er * tr[p] - ei * ti[p]) + zr * e;
            const float ni = (er * ti[p] + ei * tr[p]) + zi * e;
            tr[p] = nr;
            ti[p] = ni;
        } else {
            const float *c = D + (long long)m * 3 * P + 3 * p;
            if (P == 1) phi = c[0] * tr[p];
            else phi = phi + c[0] * tr[p];
            tr[p] = c[1] * tr[p] + c[2] * e;
        }
    }
    return phi;
}

// Dispersive part B (gprMax/fields_updates_ext.pyx:183-235, 299-346).
template <int P, bool CPLX>
__device__ __forceinline__ void disp_b(const float *D, uint32_t m, float *tr, float *ti, float e)
{
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (CPLX) {
            const float *c = D + ((long long)m * 3 * P + 3 * p) * 2;
            tr[p] = tr[p] - c[4] * e;
            ti[p] = ti[p] - c[5] * e;
        } else {
            tr[p] = tr[p] - D[(long long)m * 3 * P + 3 * p + 2] * e;
        }
    }
}

template <int P, bool CPLX>
__device__ __forceinline__ void t_load(float *const *T, int c, long long N, long long f, float *tr,
                                       float *ti)
{
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (CPLX) {
            const float2 v = reinterpret_cast<const float2 *>(T[c])[(long long)p * N + f];
            tr[p] = v.x;
            ti[p] = v.y;
        } else {
            tr[p] = T[c][(long long)p * N + f];
            ti[p] = 0.f;
        }
    }
}

template <int P, bool CPLX>
__device__ __forceinline__ void t_store(float *const *T, int c, long long N, long long f,
                                        const float *tr, const float *ti)
{
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (CPLX)
            reinterpret_cast<float2 *>(T[c])[(long long)p * N + f] = make_float2(tr[p], ti[p]);
        else
            T[c][(long long)p * N + f] = tr[p];
    }
}

template <int FORM, int P, bool CPLX>
__global__ void k_e_pass(StepArgs a)
{
    extern __shared__ float lds[];
    const int l = logical_block(a);
    if (l >= a.nblocks) return;
    const int nc = a.nmat * 5;
    const int nd = P > 0 ? a.nmat * 3 * P * (CPLX ? 2 : 1) : 0;
    float *ldisp = lds + nc;
    uint8_t *lflag = reinterpret_cast<uint8_t *>(ldisp + nd);
    for (int n = threadIdx.x; n < nc; n += blockDim.x) lds[n] = a.coef[n];
    if (P > 0) {
        for (int n = threadIdx.x; n < nd; n += blockDim.x) ldisp[n] = a.disp[n];
        for (int n = threadIdx.x; n < a.nmat; n += blockDim.x) lflag[n] = a.dispflag[n];
    }
    __syncthreads();

    const int seg = l % a.nseg, chunk = l / a.nseg;
    const int t = seg * blockDim.x + threadIdx.x;
    if (t >= a.plane) return;
    const int NZ1 = a.nz + 1;
    const int j = t / NZ1, k = t - j * NZ1;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = chunk * a.xc;
    const int i1 = min(i0 + a.xc, nx + 1);
    const long long Pl = a.plane, N = a.ntot;
    const int flags = a.flags;
    const uint32_t *ID = a.ID;

    long long f = (long long)i0 * Pl + t;
    float hy_m = 0.f, hz_m = 0.f;  // Hy, Hz of plane i-1
    if (i0 >= 1) {
        hy_m = a.Hy[f - Pl];
        hz_m = a.Hz[f - Pl];
    }
    for (int i = i0; i < i1; ++i, f += Pl) {
        const float hx = a.Hx[f], hy = a.Hy[f], hz = a.Hz[f];
        float ex = a.Ex[f], ey = a.Ey[f], ez = a.Ez[f];
        // Differences of H used by the E update and the E-PML (backward).
        const float dHz_y = (j >= 1) ? hz - a.Hz[f - NZ1] : 0.f;
        const float dHy_z = (k >= 1) ? hy - a.Hy[f - 1] : 0.f;
        const float dHx_z = (k >= 1) ? hx - a.Hx[f - 1] : 0.f;
        const float dHz_x = (i >= 1) ? hz - hz_m : 0.f;
        const float dHy_x = (i >= 1) ? hy - hy_m : 0.f;
        const float dHx_y = (j >= 1) ? hx - a.Hx[f - NZ1] : 0.f;
        const uint32_t m0 = ID[f], m1 = ID[N + f], m2 = ID[2 * N + f];

        const bool rx_ = i < nx && j >= 1 && j < ny && k >= 1 && k < nz;
        const bool ry_ = i >= 1 && i < nx && j < ny && k >= 1 && k < nz;
        const bool rz_ = i >= 1 && i < nx && j >= 1 && j < ny && k < nz;
        float txr[P > 0 ? P : 1], txi[P > 0 ? P : 1];
        float tyr[P > 0 ? P : 1], tyi[P > 0 ? P : 1];
        float tzr[P > 0 ? P : 1], tzi[P > 0 ? P : 1];
        bool dx_ = false, dy_ = false, dz_ = false;
        if (P > 0 && (flags & (F_MAIN | F_DISPB))) {
            dx_ = rx_ && lflag[m0];
            dy_ = ry_ && lflag[m1];
            dz_ = rz_ && lflag[m2];
            if (dx_) t_load<P, CPLX>(a.T, 0, N, f, txr, txi);
            if (dy_) t_load<P, CPLX>(a.T, 1, N, f, tyr, tyi);
            if (dz_) t_load<P, CPLX>(a.T, 2, N, f, tzr, tzi);
        }

        if (flags & F_MAIN) {  // gprMax/fields_updates_ext.pyx:81-107 and 146-179
            if (rx_) {
                float v = lds[m0 * 5] * ex + lds[m0 * 5 + 2] * dHz_y - lds[m0 * 5 + 3] * dHy_z;
                if (P > 0) {
                    const float phi = dx_ ? disp_a<P, CPLX>(ldisp, m0, txr, txi, ex) : 0.f;
                    v = v - lds[m0 * 5 + 4] * phi;
                }
                ex = v;
            }
            if (ry_) {
                float v = lds[m1 * 5] * ey + lds[m1 * 5 + 3] * dHx_z - lds[m1 * 5 + 1] * dHz_x;
                if (P > 0) {
                    const float phi = dy_ ? disp_a<P, CPLX>(ldisp, m1, tyr, tyi, ey) : 0.f;
                    v = v - lds[m1 * 5 + 4] * phi;
                }
                ey = v;
            }
            if (rz_) {
                float v = lds[m2 * 5] * ez + lds[m2 * 5 + 1] * dHy_x - lds[m2 * 5 + 2] * dHx_y;
                if (P > 0) {
                    const float phi = dz_ ? disp_a<P, CPLX>(ldisp, m2, tzr, tzi, ez) : 0.f;
                    v = v - lds[m2 * 5 + 4] * phi;
                }
                ez = v;
            }
        }

        if (flags & F_PML_ALL) {
            for (int s = 0; s < a.nslab; ++s) {
                if (!(flags & (F_PML0 << s))) continue;
                const SlabArgs &sl = a.slab[s];
                if (!in_box(sl, i, j, k)) continue;
                int p;
                long long q1, q2, o1, o2;
                slab_index(sl, i, j, k, p, q1, q2, o1, o2);
                if (sl.axis == 0) {
                    ey = pml_correct<FORM>(sl, p, ey, false, lds[m1 * 5 + 4], dHz_x / sl.d, sl.phi1, q1, o1);
                    ez = pml_correct<FORM>(sl, p, ez, true, lds[m2 * 5 + 4], dHy_x / sl.d, sl.phi2, q2, o2);
                } else if (sl.axis == 1) {
                    ex = pml_correct<FORM>(sl, p, ex, true, lds[m0 * 5 + 4], dHz_y / sl.d, sl.phi1, q1, o1);
                    ez = pml_correct<FORM>(sl, p, ez, false, lds[m2 * 5 + 4], dHx_y / sl.d, sl.phi2, q2, o2);
                } else {
                    ex = pml_correct<FORM>(sl, p, ex, false, lds[m0 * 5 + 4], dHy_z / sl.d, sl.phi1, q1, o1);
                    ey = pml_correct<FORM>(sl, p, ey, true, lds[m1 * 5 + 4], dHx_z / sl.d, sl.phi2, q2, o2);
                }
            }
        }

        if (flags & F_SRC) {
            // Voltage sources, then Hertzian dipoles (gprMax/model_build_run.py:458-461).
            for (int s = 0; s < a.src0.n; ++s) {  // gprMax/sources.py:81-120
                const int *in = a.src0.info + 6 * s;
                if (in[0] == i && in[1] == j && in[2] == k && a.it >= in[4] && a.it <= in[5]) {
                    const int pol = in[3] & 3;
                    const bool hard = in[3] >= 4;
                    const uint32_t m = pol == 0 ? m0 : (pol == 1 ? m1 : m2);
                    const float w = a.src0.waves[(long long)s * a.iterations + a.it];
                    float *F = pol == 0 ? &ex : (pol == 1 ? &ey : &ez);
                    if (hard) *F = -w / a.src0.fac[2 * s];
                    else *F = *F - lds[m * 5 + 4] * w * a.src0.fac[2 * s];
                }
            }
            for (int s = 0; s < a.src1.n; ++s) {  // gprMax/sources.py:163-193
                const int *in = a.src1.info + 6 * s;
                if (in[0] == i && in[1] == j && in[2] == k && a.it >= in[4] && a.it <= in[5]) {
                    const int pol = in[3];
                    const uint32_t m = pol == 0 ? m0 : (pol == 1 ? m1 : m2);
                    const float w = a.src1.waves[(long long)s * a.iterations + a.it];
                    const float v = lds[m * 5 + 4] * w * a.src1.fac[2 * s] * a.src1.fac[2 * s + 1];
                    if (pol == 0) ex = ex - v;
                    else if (pol == 1) ey = ey - v;
                    else ez = ez - v;
                }
            }
        }

        if (P > 0 && (flags & F_DISPB)) {
            if (dx_) disp_b<P, CPLX>(ldisp, m0, txr, txi, ex);
            if (dy_) disp_b<P, CPLX>(ldisp, m1, tyr, tyi, ey);
            if (dz_) disp_b<P, CPLX>(ldisp, m2, tzr, tzi, ez);
        }
        if (P > 0 && (flags & (F_MAIN | F_DISPB))) {
            if (dx_) t_store<P, CPLX>(a.T, 0, N, f, txr, txi);
            if (dy_) t_store<P, CPLX>(a.T, 1, N, f, tyr, tyi);
            if (dz_) t_store<P, CPLX>(a.T, 2, N, f, tzr, tzi);
        }
        if (flags & (F_MAIN | F_PML_ALL | F_SRC)) {
            a.Ex[f] = ex;
            a.Ey[f] = ey;
            a.Ez[f] = ez;
        }
        hy_m = hy;
        hz_m = hz;
    }
}

// ---------------------------------------------------------------- snapshots
// Cell averages of gprMax/snapshots_ext.pyx:58-71 over the slices of
// gprMax/snapshots.py:78-80; one thread per output cell.
__global__ void k_snapshot(int ny, int nz, int xs, int ys, int zs, int sdx, int sdy, int sdz,
                           int snx, int sny, int snz, const float *Ex, const float *Ey,
                           const float *Ez, const float *Hx, const float *Hy, const float *Hz,
                           float *out)
{
    const long long n = (long long)snx * sny * snz;
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int k = (int)(o % snz);
    const int j = (int)((o / snz) % sny);
    const int i = (int)(o / ((long long)snz * sny));
    const long long NY1 = ny + 1, NZ1 = nz + 1;
    auto at = [&](int a, int b, int c) {
        return ((long long)(xs + a * sdx) * NY1 + (ys + b * sdy)) * NZ1 + (zs + c * sdz);
    };
    out[o] = (Ex[at(i, j, k)] + Ex[at(i, j + 1, k)] + Ex[at(i, j, k + 1)] + Ex[at(i, j + 1, k + 1)]) / 4.0f;
    out[n + o] = (Ey[at(i, j, k)] + Ey[at(i + 1, j, k)] + Ey[at(i, j, k + 1)] + Ey[at(i + 1, j, k + 1)]) / 4.0f;
    out[2 * n + o] = (Ez[at(i, j, k)] + Ez[at(i + 1, j, k)] + Ez[at(i, j + 1, k)] + Ez[at(i + 1, j + 1, k)]) / 4.0f;
    out[3 * n + o] = (Hx[at(i, j, k)] + Hx[at(i + 1, j, k)]) / 2.0f;
    out[4 * n + o] = (Hy[at(i, j, k)] + Hy[at(i, j + 1, k)]) / 2.0f;
    out[5 * n + o] = (Hz[at(i, j, k)] + Hz[at(i, j, k + 1)]) / 2.0f;
}

}  // namespace

// ====================================================================== ctx
struct Slab {
    int dir, xs, xf, ys, yf, zs, zf, order, t;
    float d;
    float *phi[4] = {nullptr, nullptr, nullptr, nullptr};  // EPhi1, EPhi2, HPhi1, HPhi2
    long long dims[4][3];
    float *eR = nullptr, *hR = nullptr;
};

struct Snap {
    int xs, ys, zs, dx, dy, dz, snx, sny, snz, time;
    float *out = nullptr;  // [6][snx][sny][snz]
};

struct SrcSet {
    int n = 0;
    int *info = nullptr;
    float *fac = nullptr, *waves = nullptr;
};

struct gpx_ctx {
    gpx_desc desc;
    long long ntot = 0;
    int plane = 0;
    hipStream_t stream = nullptr;
    float *F[6] = {nullptr};  // Ex, Ey, Ez, Hx, Hy, Hz
    uint32_t *ID = nullptr;
    float *cE = nullptr, *cH = nullptr;
    int disp_complex = 0;
    float *disp = nullptr;    // real [nmat][3P] or complex [nmat][3P][2]
    uint8_t *dispflag = nullptr;
    float *T[3] = {nullptr, nullptr, nullptr};
    bool have_id = false, have_coeffs = false, have_disp = false;
    std::vector<Slab> slabs;
    SrcSet src[3];
    int nrx = 0;
    int *rxc = nullptr;
    float *rxbuf = nullptr;
    std::vector<Snap> snaps;
    int xchunk = 0, block = 0;
    size_t bytes = 0;
    std::string err;
};

static std::string g_err;

#define GPX_TRY(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail_hip(ctx, e_, #call);                                 \
    } while (0)

static int fail(gpx_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    else g_err = msg;
    return code;
}

static int fail_hip(gpx_ctx *ctx, hipError_t e, const char *what)
{
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, (int)e, msg);
}

static hipError_t dalloc(gpx_ctx *ctx, void **p, size_t bytes, bool zero = true)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) return e;
    ctx->bytes += bytes;
    if (zero) e = hipMemsetAsync(*p, 0, bytes ? bytes : 4, ctx->stream);
    return e;
}

extern "C" const char *gpx_last_error(const gpx_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_err.c_str();
}

extern "C" int gpx_create(const gpx_desc *d, gpx_ctx **out)
{
    gpx_ctx *ctx = nullptr;
    if (!d || !out) return fail(nullptr, GPX_EARG, "gpx_create: null argument");
    *out = nullptr;
    if (d->nx < 2 || d->ny < 2 || d->nz < 2)
        return fail(nullptr, GPX_EARG, "gpx_create: only 3D grids (nx, ny, nz >= 2) are supported");
    if (d->iterations < 1 || d->nmat < 1)
        return fail(nullptr, GPX_EARG, "gpx_create: iterations and nmat must be positive");
    if (d->maxpoles < 0 || d->maxpoles > MAX_POLES)
        return fail(nullptr, GPX_ELIMIT, "gpx_create: at most " + std::to_string(MAX_POLES) + " dispersive poles are supported");
    if (d->pml_formulation != 0 && d->pml_formulation != 1)
        return fail(nullptr, GPX_EARG, "gpx_create: pml_formulation must be 0 (HORIPML) or 1 (MRIPML)");
    size_t lds = (size_t)d->nmat * 5 * 4 + (size_t)d->nmat * 3 * d->maxpoles * 2 * 4 + d->nmat;
    if (lds > (size_t)MAX_LDS_BYTES)
        return fail(nullptr, GPX_ELIMIT, "gpx_create: too many materials for the LDS coefficient table");
    hipError_t e = hipSetDevice(d->device);
    if (e != hipSuccess) return fail(nullptr, (int)e, std::string("hipSetDevice: ") + hipGetErrorString(e));
    ctx = new gpx_ctx();
    ctx->desc = *d;
    ctx->plane = (d->ny + 1) * (d->nz + 1);
    ctx->ntot = (long long)(d->nx + 1) * ctx->plane;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        int rc = fail(nullptr, (int)e, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        delete ctx;
        return rc;
    }
    const size_t fb = (size_t)ctx->ntot * 4;
    for (int c = 0; c < 6; ++c) {
        e = dalloc(ctx, (void **)&ctx->F[c], fb);
        if (e != hipSuccess) goto fail_alloc;
    }
    e = dalloc(ctx, (void **)&ctx->ID, 6 * fb, false);
    if (e != hipSuccess) goto fail_alloc;
    e = dalloc(ctx, (void **)&ctx->cE, (size_t)d->nmat * 5 * 4);
    if (e != hipSuccess) goto fail_alloc;
    e = dalloc(ctx, (void **)&ctx->cH, (size_t)d->nmat * 5 * 4);
    if (e != hipSuccess) goto fail_alloc;
    *out = ctx;
    return 0;
fail_alloc : {
    int rc = fail(nullptr, (int)e, std::string("hipMalloc: ") + hipGetErrorString(e));
    gpx_destroy(ctx);
    return rc;
}
}

extern "C" int gpx_destroy(gpx_ctx *ctx)
{
    if (!ctx) return 0;
    hipSetDevice(ctx->desc.device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->F) if (p) hipFree(p);
    for (auto &p : ctx->T) if (p) hipFree(p);
    if (ctx->ID) hipFree(ctx->ID);
    if (ctx->cE) hipFree(ctx->cE);
    if (ctx->cH) hipFree(ctx->cH);
    if (ctx->disp) hipFree(ctx->disp);
    if (ctx->dispflag) hipFree(ctx->dispflag);
    for (auto &s : ctx->slabs) {
        for (auto &p : s.phi) if (p) hipFree(p);
        if (s.eR) hipFree(s.eR);
        if (s.hR) hipFree(s.hR);
    }
    for (auto &s : ctx->src) {
        if (s.info) hipFree(s.info);
        if (s.fac) hipFree(s.fac);
        if (s.waves) hipFree(s.waves);
    }
    if (ctx->rxc) hipFree(ctx->rxc);
    if (ctx->rxbuf) hipFree(ctx->rxbuf);
    for (auto &s : ctx->snaps) if (s.out) hipFree(s.out);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

static int copy_in(gpx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    GPX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    GPX_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

static int copy_out(gpx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    GPX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    GPX_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int gpx_upload_id(gpx_ctx *ctx, const uint32_t *ID)
{
    if (!ctx || !ID) return fail(ctx, GPX_EARG, "gpx_upload_id: null argument");
    const uint32_t nmat = (uint32_t)ctx->desc.nmat;
    const size_t n = 6 * (size_t)ctx->ntot;
    // Every ID indexes the coefficient tables; reject out-of-range values on the
    // host so no kernel can read outside them.
    for (size_t q = 0; q < n; ++q)
        if (ID[q] >= nmat)
            return fail(ctx, GPX_EARG, "gpx_upload_id: material index " + std::to_string(ID[q]) +
                                           " >= nmat " + std::to_string(nmat));
    hipSetDevice(ctx->desc.device);
    int rc = copy_in(ctx, ctx->ID, ID, n * 4);
    if (!rc) ctx->have_id = true;
    return rc;
}

extern "C" int gpx_upload_coeffs(gpx_ctx *ctx, const float *cE, const float *cH)
{
    if (!ctx || !cE || !cH) return fail(ctx, GPX_EARG, "gpx_upload_coeffs: null argument");
    hipSetDevice(ctx->desc.device);
    const size_t b = (size_t)ctx->desc.nmat * 5 * 4;
    int rc = copy_in(ctx, ctx->cE, cE, b);
    if (!rc) rc = copy_in(ctx, ctx->cH, cH, b);
    if (!rc) ctx->have_coeffs = true;
    return rc;
}

extern "C" int gpx_upload_dispersive(gpx_ctx *ctx, const float *dc)
{
    if (!ctx || !dc) return fail(ctx, GPX_EARG, "gpx_upload_dispersive: null argument");
    const int P = ctx->desc.maxpoles, nm = ctx->desc.nmat;
    if (P == 0) return fail(ctx, GPX_ESTATE, "gpx_upload_dispersive: model has no poles");
    hipSetDevice(ctx->desc.device);
    // Keep complex arithmetic only where an imaginary part is non-zero
    // (Lorentz); Debye and Drude poles are real.
    int cplx = 0;
    for (size_t q = 0; q < (size_t)nm * 3 * P; ++q)
        if (dc[2 * q + 1] != 0.0f) cplx = 1;
    std::vector<float> host;
    std::vector<uint8_t> flag(nm, 0);
    for (int m = 0; m < nm; ++m)
        for (int q = 0; q < 3 * P; ++q)
            if (dc[((size_t)m * 3 * P + q) * 2] != 0.0f || dc[((size_t)m * 3 * P + q) * 2 + 1] != 0.0f) flag[m] = 1;
    if (cplx) host.assign(dc, dc + (size_t)nm * 3 * P * 2);
    else {
        host.resize((size_t)nm * 3 * P);
        for (size_t q = 0; q < host.size(); ++q) host[q] = dc[2 * q];
    }
    if (ctx->have_disp && cplx != ctx->disp_complex)
        return fail(ctx, GPX_ESTATE, "gpx_upload_dispersive: real/complex kind cannot change after upload");
    if (!ctx->disp) {
        GPX_TRY(dalloc(ctx, (void **)&ctx->disp, host.size() * 4));
        GPX_TRY(dalloc(ctx, (void **)&ctx->dispflag, nm));
        const size_t tb = (size_t)P * ctx->ntot * 4 * (cplx ? 2 : 1);
        for (int c = 0; c < 3; ++c) GPX_TRY(dalloc(ctx, (void **)&ctx->T[c], tb));
    }
    ctx->disp_complex = cplx;
    int rc = copy_in(ctx, ctx->disp, host.data(), host.size() * 4);
    if (!rc) rc = copy_in(ctx, ctx->dispflag, flag.data(), nm);
    if (!rc) ctx->have_disp = true;
    return rc;
}

static void phi_dims(int field, int axis, int snx, int sny, int snz, long long d1[3], long long d2[3])
{
    // gprMax/pml.py:202-219
    long long a[3], b[3];
    if (field == 0) {
        if (axis == 0) { a[0] = snx + 1; a[1] = sny; a[2] = snz + 1; b[0] = snx + 1; b[1] = sny + 1; b[2] = snz; }
        else if (axis == 1) { a[0] = snx; a[1] = sny + 1; a[2] = snz + 1; b[0] = snx + 1; b[1] = sny + 1; b[2] = snz; }
        else { a[0] = snx; a[1] = sny + 1; a[2] = snz + 1; b[0] = snx + 1; b[1] = sny; b[2] = snz + 1; }
    } else {
        if (axis == 0) { a[0] = snx; a[1] = sny + 1; a[2] = snz; b[0] = snx; b[1] = sny; b[2] = snz + 1; }
        else if (axis == 1) { a[0] = snx + 1; a[1] = sny; a[2] = snz; b[0] = snx; b[1] = sny; b[2] = snz + 1; }
        else { a[0] = snx + 1; a[1] = sny; a[2] = snz; b[0] = snx; b[1] = sny + 1; b[2] = snz; }
    }
    for (int n = 0; n < 3; ++n) { d1[n] = a[n]; d2[n] = b[n]; }
}

extern "C" int gpx_add_pml(gpx_ctx *ctx, int dir, int xs, int xf, int ys, int yf, int zs, int zf,
                           int order, float d, const float *eR, const float *hR, int *slab)
{
    if (!ctx || !eR || !hR) return fail(ctx, GPX_EARG, "gpx_add_pml: null argument");
    if ((int)ctx->slabs.size() >= MAX_SLABS) return fail(ctx, GPX_ELIMIT, "gpx_add_pml: at most 6 slabs");
    if (dir < 0 || dir > 5 || (order != 1 && order != 2))
        return fail(ctx, GPX_EARG, "gpx_add_pml: direction must be 0..5 and order 1 or 2");
    const gpx_desc &g = ctx->desc;
    if (xs < 0 || ys < 0 || zs < 0 || xf > g.nx || yf > g.ny || zf > g.nz || xs >= xf || ys >= yf || zs >= zf)
        return fail(ctx, GPX_EARG, "gpx_add_pml: slab box outside the grid");
    const int axis = dir % 3;
    const int t = axis == 0 ? xf - xs : (axis == 1 ? yf - ys : zf - zs);
    // The corrected cells reach one cell beyond the box on the minus side
    // (E) and one cell inside it (H); both must have their neighbours.
    const int lo = axis == 0 ? xs : (axis == 1 ? ys : zs);
    const int hi = axis == 0 ? xf : (axis == 1 ? yf : zf);
    const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
    if (dir < 3 && (hi - t + 1 < 1 || hi >= n + 1)) return fail(ctx, GPX_EARG, "gpx_add_pml: minus slab out of range");
    if (dir >= 3 && (lo < 1 || lo + t > n)) return fail(ctx, GPX_EARG, "gpx_add_pml: plus slab out of range");
    hipSetDevice(g.device);
    Slab s;
    s.dir = dir; s.xs = xs; s.xf = xf; s.ys = ys; s.yf = yf; s.zs = zs; s.zf = zf;
    s.order = order; s.t = t; s.d = d;
    phi_dims(0, axis, xf - xs, yf - ys, zf - zs, s.dims[0], s.dims[1]);
    phi_dims(1, axis, xf - xs, yf - ys, zf - zs, s.dims[2], s.dims[3]);
    for (int q = 0; q < 4; ++q) {
        size_t b = (size_t)order * s.dims[q][0] * s.dims[q][1] * s.dims[q][2] * 4;
        GPX_TRY(dalloc(ctx, (void **)&s.phi[q], b));
    }
    const size_t rb = (size_t)4 * order * t * 4;
    GPX_TRY(dalloc(ctx, (void **)&s.eR, rb, false));
    GPX_TRY(dalloc(ctx, (void **)&s.hR, rb, false));
    ctx->slabs.push_back(s);
    Slab &ref = ctx->slabs.back();
    int rc = copy_in(ctx, ref.eR, eR, rb);
    if (!rc) rc = copy_in(ctx, ref.hR, hR, rb);
    if (slab) *slab = (int)ctx->slabs.size() - 1;
    return rc;
}

extern "C" int gpx_set_sources(gpx_ctx *ctx, int kind, int n, const int32_t *info,
                               const float *factors, const float *waves)
{
    if (!ctx || kind < 0 || kind > 2 || n < 0) return fail(ctx, GPX_EARG, "gpx_set_sources: bad argument");
    if (n > 0 && (!info || !factors || !waves)) return fail(ctx, GPX_EARG, "gpx_set_sources: null array");
    const gpx_desc &g = ctx->desc;
    for (int s = 0; s < n; ++s) {
        const int32_t *in = info + 6 * s;
        const int pol = in[3] & 3;
        if (in[0] < 0 || in[0] > g.nx || in[1] < 0 || in[1] > g.ny || in[2] < 0 || in[2] > g.nz || pol > 2 ||
            (in[3] >= 4 && kind != GPX_SRC_VOLTAGE) || in[3] < 0 || in[3] > 6)
            return fail(ctx, GPX_EARG, "gpx_set_sources: source " + std::to_string(s) + " has a bad cell or polarisation");
    }
    hipSetDevice(g.device);
    SrcSet &S = ctx->src[kind];
    if (S.info) { hipFree(S.info); S.info = nullptr; }
    if (S.fac) { hipFree(S.fac); S.fac = nullptr; }
    if (S.waves) { hipFree(S.waves); S.waves = nullptr; }
    S.n = n;
    if (n == 0) return 0;
    GPX_TRY(dalloc(ctx, (void **)&S.info, (size_t)n * 6 * 4, false));
    GPX_TRY(dalloc(ctx, (void **)&S.fac, (size_t)n * 2 * 4, false));
    GPX_TRY(dalloc(ctx, (void **)&S.waves, (size_t)n * g.iterations * 4, false));
    int rc = copy_in(ctx, S.info, info, (size_t)n * 6 * 4);
    if (!rc) rc = copy_in(ctx, S.fac, factors, (size_t)n * 2 * 4);
    if (!rc) rc = copy_in(ctx, S.waves, waves, (size_t)n * g.iterations * 4);
    return rc;
}

extern "C" int gpx_set_receivers(gpx_ctx *ctx, int n, const int32_t *coords)
{
    if (!ctx || n < 0 || (n > 0 && !coords)) return fail(ctx, GPX_EARG, "gpx_set_receivers: bad argument");
    const gpx_desc &g = ctx->desc;
    for (int r = 0; r < n; ++r)
        if (coords[3 * r] < 0 || coords[3 * r] > g.nx || coords[3 * r + 1] < 0 || coords[3 * r + 1] > g.ny ||
            coords[3 * r + 2] < 0 || coords[3 * r + 2] > g.nz)
            return fail(ctx, GPX_EARG, "gpx_set_receivers: receiver outside the grid");
    hipSetDevice(g.device);
    if (ctx->rxc) { hipFree(ctx->rxc); ctx->rxc = nullptr; }
    if (ctx->rxbuf) { hipFree(ctx->rxbuf); ctx->rxbuf = nullptr; }
    ctx->nrx = n;
    if (n == 0) return 0;
    GPX_TRY(dalloc(ctx, (void **)&ctx->rxc, (size_t)n * 3 * 4, false));
    GPX_TRY(dalloc(ctx, (void **)&ctx->rxbuf, (size_t)6 * g.iterations * n * 4));
    return copy_in(ctx, ctx->rxc, coords, (size_t)n * 3 * 4);
}

extern "C" int gpx_add_snapshot(gpx_ctx *ctx, int xs, int ys, int zs, int dx, int dy, int dz,
                                int snx, int sny, int snz, int time, int *index)
{
    if (!ctx || dx < 1 || dy < 1 || dz < 1 || snx < 1 || sny < 1 || snz < 1 || xs < 0 || ys < 0 || zs < 0)
        return fail(ctx, GPX_EARG, "gpx_add_snapshot: bad argument");
    const gpx_desc &g = ctx->desc;
    // The averages read slice points up to index snx (one past the last cell).
    if (xs + (long long)snx * dx > g.nx || ys + (long long)sny * dy > g.ny || zs + (long long)snz * dz > g.nz)
        return fail(ctx, GPX_EARG, "gpx_add_snapshot: snapshot volume outside the grid");
    hipSetDevice(g.device);
    Snap s;
    s.xs = xs; s.ys = ys; s.zs = zs; s.dx = dx; s.dy = dy; s.dz = dz;
    s.snx = snx; s.sny = sny; s.snz = snz; s.time = time;
    GPX_TRY(dalloc(ctx, (void **)&s.out, (size_t)6 * snx * sny * snz * 4));
    ctx->snaps.push_back(s);
    if (index) *index = (int)ctx->snaps.size() - 1;
    return 0;
}

extern "C" int gpx_set_launch(gpx_ctx *ctx, int xchunk, int block)
{
    if (!ctx || xchunk < 0 || !(block == 0 || block == 64 || block == 128 || block == 256 || block == 512 || block == 1024))
        return fail(ctx, GPX_EARG, "gpx_set_launch: bad argument");
    ctx->xchunk = xchunk;
    ctx->block = block;
    return 0;
}

// ---------------------------------------------------------------- launching
static void launch_shape(const gpx_ctx *ctx, int &block, int &xc, int &nseg, int &nchunk)
{
    block = ctx->block ? ctx->block : 256;
    nseg = (ctx->plane + block - 1) / block;
    const int NX1 = ctx->desc.nx + 1;
    if (ctx->xchunk) xc = ctx->xchunk;
    else {
        // Aim for about 8 blocks per CU of the 256 CUs, at least 8 planes per chunk.
        long long want = 2048;
        long long chunks = (want + nseg - 1) / nseg;
        xc = (int)((NX1 + chunks - 1) / chunks);
        if (xc < 8) xc = 8;
    }
    if (xc > NX1) xc = NX1;
    nchunk = (NX1 + xc - 1) / xc;
}

static void fill_slab_args(const gpx_ctx *ctx, int field, int s, SlabArgs &a)
{
    const Slab &S = ctx->slabs[s];
    const int axis = S.dir % 3, minus = S.dir < 3;
    const int st[3] = {S.xs, S.ys, S.zs}, en[3] = {S.xf, S.yf, S.zf};
    a.axis = axis;
    a.minus = minus;
    for (int n = 0; n < 3; ++n) {
        a.lo[n] = st[n];
        a.hi[n] = en[n];
        a.org[n] = st[n];
    }
    if (minus) {
        if (field == 0) {  // cells f - p, p in [0, t)
            a.lo[axis] = en[axis] - S.t + 1;
            a.hi[axis] = en[axis] + 1;
            a.pend = en[axis];
        } else {           // cells f - (p + 1)
            a.lo[axis] = en[axis] - S.t;
            a.hi[axis] = en[axis];
            a.pend = en[axis] - 1;
        }
    } else {
        a.lo[axis] = st[axis];
        a.hi[axis] = st[axis] + S.t;
        a.pend = st[axis];
    }
    const int q = field == 0 ? 0 : 2;
    for (int n = 0; n < 3; ++n) {
        a.d1[n] = (int)S.dims[q][n];
        a.d2[n] = (int)S.dims[q + 1][n];
    }
    a.t = S.t;
    a.order = S.order;
    a.d = S.d;
    a.phi1 = S.phi[q];
    a.phi2 = S.phi[q + 1];
    a.R = field == 0 ? S.eR : S.hR;
}

static void base_args(const gpx_ctx *ctx, int field, StepArgs &a, int &block, int &grid, size_t &lds)
{
    int xc, nseg, nchunk;
    launch_shape(ctx, block, xc, nseg, nchunk);
    memset(&a, 0, sizeof(a));
    const gpx_desc &g = ctx->desc;
    a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
    a.plane = ctx->plane;
    a.nseg = nseg;
    a.nblocks = nseg * nchunk;
    a.per_xcd = (a.nblocks + 7) / 8;
    a.xc = xc;
    a.Ex = ctx->F[0]; a.Ey = ctx->F[1]; a.Ez = ctx->F[2];
    a.Hx = ctx->F[3]; a.Hy = ctx->F[4]; a.Hz = ctx->F[5];
    a.ID = ctx->ID;
    a.ntot = ctx->ntot;
    a.nmat = g.nmat;
    a.coef = field == 0 ? ctx->cE : ctx->cH;
    a.iterations = g.iterations;
    a.nslab = (int)ctx->slabs.size();
    for (int s = 0; s < a.nslab; ++s) fill_slab_args(ctx, field, s, a.slab[s]);
    lds = (size_t)g.nmat * 5 * 4;
    if (field == 0) {
        a.src0 = {ctx->src[GPX_SRC_VOLTAGE].n, ctx->src[GPX_SRC_VOLTAGE].info, ctx->src[GPX_SRC_VOLTAGE].fac, ctx->src[GPX_SRC_VOLTAGE].waves};
        a.src1 = {ctx->src[GPX_SRC_HERTZIAN].n, ctx->src[GPX_SRC_HERTZIAN].info, ctx->src[GPX_SRC_HERTZIAN].fac, ctx->src[GPX_SRC_HERTZIAN].waves};
        const int P = g.maxpoles;
        if (P > 0) {
            a.disp = ctx->disp;
            a.dispflag = ctx->dispflag;
            for (int c = 0; c < 3; ++c) a.T[c] = ctx->T[c];
            lds += (size_t)g.nmat * 3 * P * 4 * (ctx->disp_complex ? 2 : 1) + g.nmat;
        }
    } else {
        a.src0 = {ctx->src[GPX_SRC_MAGNETIC].n, ctx->src[GPX_SRC_MAGNETIC].info, ctx->src[GPX_SRC_MAGNETIC].fac, ctx->src[GPX_SRC_MAGNETIC].waves};
        a.nrx = ctx->nrx;
        a.rxc = ctx->rxc;
        a.rxbuf = ctx->rxbuf;
    }
    grid = a.per_xcd * 8;
}

template <int FORM>
static hipError_t launch_e_form(const gpx_ctx *ctx, const StepArgs &a, int grid, int block, size_t lds)
{
    const int P = ctx->desc.maxpoles;
    const bool C = ctx->disp_complex != 0;
    hipStream_t st = ctx->stream;
#define GPX_E(PP, CC) hipLaunchKernelGGL((k_e_pass<FORM, PP, CC>), dim3(grid), dim3(block), lds, st, a)
    switch (P) {
    case 0: GPX_E(0, false); break;
    case 1: if (C) GPX_E(1, true); else GPX_E(1, false); break;
    case 2: if (C) GPX_E(2, true); else GPX_E(2, false); break;
    case 3: if (C) GPX_E(3, true); else GPX_E(3, false); break;
    case 4: if (C) GPX_E(4, true); else GPX_E(4, false); break;
    default: return hipErrorInvalidValue;
    }
#undef GPX_E
    return hipGetLastError();
}

static hipError_t launch_pass(const gpx_ctx *ctx, int field, int flags, int it)
{
    StepArgs a;
    int block, grid;
    size_t lds;
    base_args(ctx, field, a, block, grid, lds);
    a.flags = flags;
    a.it = it;
    if (field == 1) {
        if (ctx->desc.pml_formulation == 0)
            hipLaunchKernelGGL((k_h_pass<0>), dim3(grid), dim3(block), lds, ctx->stream, a);
        else
            hipLaunchKernelGGL((k_h_pass<1>), dim3(grid), dim3(block), lds, ctx->stream, a);
        return hipGetLastError();
    }
    return ctx->desc.pml_formulation == 0 ? launch_e_form<0>(ctx, a, grid, block, lds)
                                          : launch_e_form<1>(ctx, a, grid, block, lds);
}

static hipError_t launch_snapshot(const gpx_ctx *ctx, int s)
{
    const Snap &S = ctx->snaps[s];
    const long long n = (long long)S.snx * S.sny * S.snz;
    const int block = 256;
    const int grid = (int)((n + block - 1) / block);
    hipLaunchKernelGGL(k_snapshot, dim3(grid), dim3(block), 0, ctx->stream, ctx->desc.ny, ctx->desc.nz,
                       S.xs, S.ys, S.zs, S.dx, S.dy, S.dz, S.snx, S.sny, S.snz, ctx->F[0], ctx->F[1],
                       ctx->F[2], ctx->F[3], ctx->F[4], ctx->F[5], S.out);
    return hipGetLastError();
}

static int check_ready(gpx_ctx *ctx)
{
    if (!ctx->have_id || !ctx->have_coeffs)
        return fail(ctx, GPX_ESTATE, "ID and coefficient tables must be uploaded before stepping");
    if (ctx->desc.maxpoles > 0 && !ctx->have_disp)
        return fail(ctx, GPX_ESTATE, "dispersive coefficients must be uploaded before stepping");
    return 0;
}

static int h_flags(const gpx_ctx *ctx)
{
    int f = F_MAIN | F_SRC | (ctx->nrx ? F_RX : 0);
    for (size_t s = 0; s < ctx->slabs.size(); ++s) f |= F_PML0 << s;
    return f;
}

static int e_flags(const gpx_ctx *ctx)
{
    int f = F_MAIN | F_SRC | (ctx->desc.maxpoles ? F_DISPB : 0);
    for (size_t s = 0; s < ctx->slabs.size(); ++s) f |= F_PML0 << s;
    return f;
}

static int run_steps(gpx_ctx *ctx, int it0, int n, int per_kernel, float *ms_h, float *ms_e)
{
    const int hf = h_flags(ctx), ef = e_flags(ctx);
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    double th = 0, te = 0;
    if (per_kernel)
        for (auto &e : ev) GPX_TRY(hipEventCreate(&e));
    for (int it = it0; it < it0 + n; ++it) {
        for (size_t s = 0; s < ctx->snaps.size(); ++s)
            if (ctx->snaps[s].time == it + 1) GPX_TRY(launch_snapshot(ctx, (int)s));
        if (per_kernel) GPX_TRY(hipEventRecord(ev[0], ctx->stream));
        GPX_TRY(launch_pass(ctx, 1, hf, it));
        if (per_kernel) GPX_TRY(hipEventRecord(ev[1], ctx->stream));
        GPX_TRY(launch_pass(ctx, 0, ef, it));
        if (per_kernel) {
            GPX_TRY(hipEventRecord(ev[2], ctx->stream));
            GPX_TRY(hipEventSynchronize(ev[2]));
            float a = 0, b = 0;
            GPX_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
            GPX_TRY(hipEventElapsedTime(&b, ev[1], ev[2]));
            th += a;
            te += b;
        }
    }
    if (per_kernel) {
        for (auto &e : ev) hipEventDestroy(e);
        if (ms_h) *ms_h = (float)(th / (n > 0 ? n : 1));
        if (ms_e) *ms_e = (float)(te / (n > 0 ? n : 1));
    }
    return 0;
}

extern "C" int gpx_run(gpx_ctx *ctx, int it0, int n)
{
    if (!ctx || it0 < 0 || n < 0 || it0 + n > ctx->desc.iterations)
        return fail(ctx, GPX_EARG, "gpx_run: iteration range outside [0, iterations)");
    if (int rc = check_ready(ctx)) return rc;
    hipSetDevice(ctx->desc.device);
    return run_steps(ctx, it0, n, 0, nullptr, nullptr);
}

extern "C" int gpx_run_timed(gpx_ctx *ctx, int it0, int n, int per_kernel, float *ms_total,
                             float *ms_h, float *ms_e)
{
    if (!ctx || it0 < 0 || n < 0 || it0 + n > ctx->desc.iterations)
        return fail(ctx, GPX_EARG, "gpx_run_timed: iteration range outside [0, iterations)");
    if (int rc = check_ready(ctx)) return rc;
    hipSetDevice(ctx->desc.device);
    hipEvent_t a, b;
    GPX_TRY(hipEventCreate(&a));
    GPX_TRY(hipEventCreate(&b));
    GPX_TRY(hipEventRecord(a, ctx->stream));
    int rc = run_steps(ctx, it0, n, per_kernel, ms_h, ms_e);
    if (rc) return rc;
    GPX_TRY(hipEventRecord(b, ctx->stream));
    GPX_TRY(hipEventSynchronize(b));
    float ms = 0;
    GPX_TRY(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (ms_total) *ms_total = ms;
    return 0;
}

extern "C" int gpx_part(gpx_ctx *ctx, int part, int arg, int it)
{
    if (!ctx) return fail(ctx, GPX_EARG, "gpx_part: null ctx");
    if (it < 0 || it >= ctx->desc.iterations) return fail(ctx, GPX_EARG, "gpx_part: iteration out of range");
    if (int rc = check_ready(ctx)) return rc;
    hipSetDevice(ctx->desc.device);
    const int ns = (int)ctx->slabs.size();
    switch (part) {
    case GPX_PART_STORE_RX:
        if (ctx->nrx) GPX_TRY(launch_pass(ctx, 1, F_RX, it));
        return 0;
    case GPX_PART_SNAPSHOT:
        if (arg < 0 || arg >= (int)ctx->snaps.size()) return fail(ctx, GPX_EARG, "gpx_part: bad snapshot index");
        GPX_TRY(launch_snapshot(ctx, arg));
        return 0;
    case GPX_PART_H_MAIN: GPX_TRY(launch_pass(ctx, 1, F_MAIN, it)); return 0;
    case GPX_PART_H_SRC: GPX_TRY(launch_pass(ctx, 1, F_SRC, it)); return 0;
    case GPX_PART_E_MAIN: GPX_TRY(launch_pass(ctx, 0, F_MAIN, it)); return 0;
    case GPX_PART_E_SRC: GPX_TRY(launch_pass(ctx, 0, F_SRC, it)); return 0;
    case GPX_PART_E_DISP_B:
        if (ctx->desc.maxpoles) GPX_TRY(launch_pass(ctx, 0, F_DISPB, it));
        return 0;
    case GPX_PART_H_PML:
    case GPX_PART_E_PML:
        if (arg < 0 || arg >= ns) return fail(ctx, GPX_EARG, "gpx_part: bad slab index");
        GPX_TRY(launch_pass(ctx, part == GPX_PART_H_PML ? 1 : 0, F_PML0 << arg, it));
        return 0;
    default: return fail(ctx, GPX_EARG, "gpx_part: unknown part");
    }
}

extern "C" int gpx_synchronize(gpx_ctx *ctx)
{
    if (!ctx) return fail(ctx, GPX_EARG, "gpx_synchronize: null ctx");
    GPX_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int gpx_download_rx(gpx_ctx *ctx, float *out)
{
    if (!ctx || !out) return fail(ctx, GPX_EARG, "gpx_download_rx: null argument");
    if (!ctx->nrx) return 0;
    return copy_out(ctx, out, ctx->rxbuf, (size_t)6 * ctx->desc.iterations * ctx->nrx * 4);
}

extern "C" int gpx_download_fields(gpx_ctx *ctx, float *out6)
{
    if (!ctx || !out6) return fail(ctx, GPX_EARG, "gpx_download_fields: null argument");
    for (int c = 0; c < 6; ++c)
        if (int rc = copy_out(ctx, out6 + (size_t)c * ctx->ntot, ctx->F[c], (size_t)ctx->ntot * 4)) return rc;
    return 0;
}

extern "C" int gpx_upload_fields(gpx_ctx *ctx, const float *in6)
{
    if (!ctx || !in6) return fail(ctx, GPX_EARG, "gpx_upload_fields: null argument");
    for (int c = 0; c < 6; ++c)
        if (int rc = copy_in(ctx, ctx->F[c], in6 + (size_t)c * ctx->ntot, (size_t)ctx->ntot * 4)) return rc;
    return 0;
}

static int phi_ref(gpx_ctx *ctx, int slab, int which, float **p, size_t *bytes)
{
    if (slab < 0 || slab >= (int)ctx->slabs.size() || which < 0 || which > 3)
        return fail(ctx, GPX_EARG, "bad slab or Phi index");
    const Slab &s = ctx->slabs[slab];
    *p = s.phi[which];
    *bytes = (size_t)s.order * s.dims[which][0] * s.dims[which][1] * s.dims[which][2] * 4;
    return 0;
}

extern "C" int gpx_download_pml_phi(gpx_ctx *ctx, int slab, int which, float *out)
{
    if (!ctx || !out) return fail(ctx, GPX_EARG, "gpx_download_pml_phi: null argument");
    float *p;
    size_t b;
    if (int rc = phi_ref(ctx, slab, which, &p, &b)) return rc;
    return copy_out(ctx, out, p, b);
}

extern "C" int gpx_upload_pml_phi(gpx_ctx *ctx, int slab, int which, const float *in)
{
    if (!ctx || !in) return fail(ctx, GPX_EARG, "gpx_upload_pml_phi: null argument");
    float *p;
    size_t b;
    if (int rc = phi_ref(ctx, slab, which, &p, &b)) return rc;
    return copy_in(ctx, p, in, b);
}

extern "C" int gpx_download_dispersive_T(gpx_ctx *ctx, float *out)
{
    if (!ctx || !out) return fail(ctx, GPX_EARG, "gpx_download_dispersive_T: null argument");
    const int P = ctx->desc.maxpoles;
    if (!P || !ctx->T[0]) return fail(ctx, GPX_ESTATE, "gpx_download_dispersive_T: no dispersive state");
    const size_t n = (size_t)P * ctx->ntot;
    if (ctx->disp_complex) {
        for (int c = 0; c < 3; ++c)
            if (int rc = copy_out(ctx, out + (size_t)c * n * 2, ctx->T[c], n * 8)) return rc;
        return 0;
    }
    std::vector<float> tmp(n);
    for (int c = 0; c < 3; ++c) {
        if (int rc = copy_out(ctx, tmp.data(), ctx->T[c], n * 4)) return rc;
        float *o = out + (size_t)c * n * 2;
        for (size_t q = 0; q < n; ++q) {
            o[2 * q] = tmp[q];
            o[2 * q + 1] = 0.0f;
        }
    }
    return 0;
}

extern "C" int gpx_upload_dispersive_T(gpx_ctx *ctx, const float *in)
{
    if (!ctx || !in) return fail(ctx, GPX_EARG, "gpx_upload_dispersive_T: null argument");
    const int P = ctx->desc.maxpoles;
    if (!P || !ctx->T[0]) return fail(ctx, GPX_ESTATE, "gpx_upload_dispersive_T: no dispersive state");
    const size_t n = (size_t)P * ctx->ntot;
    if (ctx->disp_complex) {
        for (int c = 0; c < 3; ++c)
            if (int rc = copy_in(ctx, ctx->T[c], in + (size_t)c * n * 2, n * 8)) return rc;
        return 0;
    }
    std::vector<float> tmp(n);
    for (int c = 0; c < 3; ++c) {
        const float *s = in + (size_t)c * n * 2;
        for (size_t q = 0; q < n; ++q) {
            if (s[2 * q + 1] != 0.0f)
                return fail(ctx, GPX_EARG, "gpx_upload_dispersive_T: imaginary part in a real-pole model");
            tmp[q] = s[2 * q];
        }
        if (int rc = copy_in(ctx, ctx->T[c], tmp.data(), n * 4)) return rc;
    }
    return 0;
}

extern "C" int gpx_download_snapshot(gpx_ctx *ctx, int index, float *out6)
{
    if (!ctx || !out6 || index < 0 || index >= (int)ctx->snaps.size())
        return fail(ctx, GPX_EARG, "gpx_download_snapshot: bad argument");
    const Snap &s = ctx->snaps[index];
    return copy_out(ctx, out6, s.out, (size_t)6 * s.snx * s.sny * s.snz * 4);
}

extern "C" int gpx_device_info(int device, char *name, int namelen, size_t *totalmem,
                               size_t *freemem, int *cus)
{
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, (int)e, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    if (name && namelen > 0) {
        snprintf(name, (size_t)namelen, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (totalmem) *totalmem = prop.totalGlobalMem;
    if (cus) *cus = prop.multiProcessorCount;
    if (freemem) {
        size_t fr = 0, tot = 0;
        e = hipSetDevice(device);
        if (e == hipSuccess) e = hipMemGetInfo(&fr, &tot);
        if (e != hipSuccess) return fail(nullptr, (int)e, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
        *freemem = fr;
    }
    return 0;
}

extern "C" size_t gpx_memory_bytes(const gpx_ctx *ctx)
{
    return ctx ? ctx->bytes : 0;
}
