// Device code shared by the translation units of libgprmax_hip.so: launch
// argument types, buffer-resource helpers, the PML, dispersive and table
// helpers, and the fused time-step kernel template k_step (instantiated per
// PML formulation and order in step_*.hip).
#ifndef GPX_STEP_H
#define GPX_STEP_H

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include <type_traits>

namespace gpx {

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
    int roff;        // offset of this slab's [4][order][t] table in the LDS R area
    float d;
    float *phi1, *phi2;   // Phi read
    float *phw1, *phw2;   // Phi written (the same arrays except in the fused step)
};

struct SrcArgs {
    int n;
    const int *info;     // [n][6]
    const float *fac;    // [n][2]
    const float *waves;  // [n][iterations]
};

constexpr uint32_t PF_RX = 1, PF_SRCH = 2, PF_SRCE = 4;
constexpr int PF_SLAB_E = 8, PF_SLAB_H = 16;

// Streamed arrays are read and written with buffer instructions: a scalar
// resource per array (base address in SGPRs), the lane's byte offset within
// the plane in a VGPR and the plane's byte offset in an SGPR, so no
// per-access vector address arithmetic is needed.  Offsets are 32-bit: an
// array holds fewer than 2^30 floats.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t rsrc(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, -1, 0x00020000);
}
// The b32 buffer builtins move raw 32-bit words; floats pass through as bits.
__device__ __forceinline__ float bld(rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(rsrc_t r, uint32_t voff, uint32_t soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
template <class IDT>
__device__ __forceinline__ uint32_t bld_id(rsrc_t r, uint32_t voff, uint32_t soff)
{
    if (sizeof(IDT) == 1) return __builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0);
    if (sizeof(IDT) == 2) return __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}

// Loads and stores at a uniform base pointer plus a 32-bit lane byte offset.
template <class T>
__device__ __forceinline__ T ld(const T *base, uint32_t boff)
{
    // Device memory: global, not flat, accesses (the pointer may have been
    // read from memory, which hides its address space from the compiler).
    // The value moves as a 4- or 8-byte word.
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte values");
    typedef typename std::conditional<sizeof(T) == 4, uint32_t, unsigned long long>::type W;
    typedef __attribute__((address_space(1))) const W gW;
    const W w = *(gW *)(reinterpret_cast<const char *>(base) + boff);
    T v;
    __builtin_memcpy(&v, &w, sizeof(T));
    return v;
}
template <class T>
__device__ __forceinline__ void st(T *base, uint32_t boff, T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte values");
    typedef typename std::conditional<sizeof(T) == 4, uint32_t, unsigned long long>::type W;
    typedef __attribute__((address_space(1))) W gW;
    W w;
    __builtin_memcpy(&w, &v, sizeof(T));
    *(gW *)(reinterpret_cast<char *>(base) + boff) = w;
}

// ---------------------------------------------------------------- PML
// One corrected component (gprMax/pml_updates/*_ext.pyx) for formulation
// FORM (0 HORIPML, 1 MRIPML) and ORDER CFS terms.  R = [RA, RB, RE, RF] each
// [order][t] in LDS; i1/off index Phi order 0 and the order-1 offset.  Phi is
// read from `pin` and, when `wr`, the advanced values are written to `pout`
// (the same array for an in-place update).
// Phi pointers come from slab descriptions staged in LDS, so the compiler
// cannot tell their address space; Phi is always device memory, and global
// loads and stores avoid the flat path (which waits on both LDS and memory
// counters).
typedef __attribute__((address_space(1))) float gfloat_t;

// Phi values p0 (order 0) and p1 (order 1, ORDER 2 only) already loaded.
template <int FORM, int ORDER>
__device__ __forceinline__ float pml_v(const SlabArgs &s, const float *R, int p, float F,
                                       bool plus_sign, float c, float dd, float p0, float p1,
                                       float *pout_, bool wr, uint32_t i1, uint32_t off)
{
    gfloat_t *pout = (gfloat_t *)pout_;
    const int ot = ORDER * s.t;
    float corr;
    if (FORM == 0) {
        if (ORDER == 1) {  // pml_updates_*_HORIPML_ext.pyx, order1_*
            const float RA01 = R[p] - 1.0f;
            const float RB0 = R[ot + p], RE0 = R[2 * ot + p], RF0 = R[3 * ot + p];
            const float ph = p0;
            corr = c * (RA01 * dd + RB0 * ph);
            if (wr) pout[i1] = RE0 * ph - RF0 * dd;
        } else {           // order2_*
            const float RA0 = R[p], RA1 = R[s.t + p];
            const float RB0 = R[ot + p], RB1 = R[ot + s.t + p];
            const float RE0 = R[2 * ot + p], RE1 = R[2 * ot + s.t + p];
            const float RF0 = R[3 * ot + p], RF1 = R[3 * ot + s.t + p];
            const float RA01 = RA0 * RA1 - 1.0f;
            corr = c * (RA01 * dd + RA1 * RB0 * p0 + RB1 * p1);
            if (wr) {
                pout[i1 + off] = RE1 * p1 - RF1 * (RA0 * dd + RB0 * p0);
                pout[i1] = RE0 * p0 - RF0 * dd;
            }
        }
    } else {
        if (ORDER == 1) {  // pml_updates_*_MRIPML_ext.pyx, order1_*
            const float IRA = 1.0f / R[p];
            const float IRA1 = IRA - 1.0f;
            const float RB0 = R[ot + p], RE0 = R[2 * ot + p], RF0 = R[3 * ot + p];
            const float RC0 = IRA * RB0 * RF0;
            const float ph = p0;
            corr = c * (IRA1 * dd - IRA * ph);
            if (wr) pout[i1] = RE0 * ph + RC0 * dd - RC0 * ph;
        } else {           // order2_*
            const float IRA = 1.0f / (R[p] + R[s.t + p]);
            const float IRA1 = IRA - 1.0f;
            const float RB0 = R[ot + p], RB1 = R[ot + s.t + p];
            const float RE0 = R[2 * ot + p], RE1 = R[2 * ot + s.t + p];
            const float RC0 = IRA * R[3 * ot + p], RC1 = IRA * R[3 * ot + s.t + p];
            const float psi = RB0 * p0 + RB1 * p1;
            corr = c * (IRA1 * dd - IRA * psi);
            if (wr) {
                pout[i1 + off] = RE1 * p1 + RC1 * (dd - psi);
                pout[i1] = RE0 * p0 + RC0 * (dd - psi);
            }
        }
    }
    return plus_sign ? F + corr : F - corr;
}

template <int FORM, int ORDER>
__device__ __forceinline__ float pml_t(const SlabArgs &s, const float *R, int p, float F,
                                       bool plus_sign, float c, float dd, const float *pin_,
                                       float *pout_, bool wr, uint32_t i1, uint32_t off)
{
    const gfloat_t *pin = (const gfloat_t *)pin_;
    const float p0 = pin[i1], p1 = ORDER == 2 ? pin[i1 + off] : 0.0f;
    return pml_v<FORM, ORDER>(s, R, p, F, plus_sign, c, dd, p0, p1, pout_, wr, i1, off);
}

// The same with formulation and order chosen at run time.
__device__ __forceinline__ float pml_apply(int form, const SlabArgs &s, const float *R, int p,
                                           float F, bool plus_sign, float c, float dd,
                                           const float *pin, float *pout, bool wr, uint32_t i1,
                                           uint32_t off)
{
    if (form == 0)
        return s.order == 1 ? pml_t<0, 1>(s, R, p, F, plus_sign, c, dd, pin, pout, wr, i1, off)
                            : pml_t<0, 2>(s, R, p, F, plus_sign, c, dd, pin, pout, wr, i1, off);
    return s.order == 1 ? pml_t<1, 1>(s, R, p, F, plus_sign, c, dd, pin, pout, wr, i1, off)
                        : pml_t<1, 2>(s, R, p, F, plus_sign, c, dd, pin, pout, wr, i1, off);
}

// One in-place PML correction with the Phi values already loaded (phi_get).
__device__ __forceinline__ float pml_correct_pre(int form, const SlabArgs &s, const float *R, int p,
                                                 float F, bool plus_sign, float c, float dd, float p0,
                                                 float p1, float *phi, uint32_t i1, uint32_t off)
{
    if (form == 0)
        return s.order == 1 ? pml_v<0, 1>(s, R, p, F, plus_sign, c, dd, p0, p1, phi, true, i1, off)
                            : pml_v<0, 2>(s, R, p, F, plus_sign, c, dd, p0, p1, phi, true, i1, off);
    return s.order == 1 ? pml_v<1, 1>(s, R, p, F, plus_sign, c, dd, p0, p1, phi, true, i1, off)
                        : pml_v<1, 2>(s, R, p, F, plus_sign, c, dd, p0, p1, phi, true, i1, off);
}

__device__ __forceinline__ bool in_box(const SlabArgs &s, int i, int j, int k)
{
    return i >= s.lo[0] && i < s.hi[0] && j >= s.lo[1] && j < s.hi[1] && k >= s.lo[2] &&
           k < s.hi[2];
}

// Profile index and the two Phi element indices of cell (i, j, k).
__device__ __forceinline__ void slab_index(const SlabArgs &s, int i, int j, int k, int &p,
                                           uint32_t &i1, uint32_t &i2, uint32_t &o1, uint32_t &o2)
{
    int c[3] = {i - s.org[0], j - s.org[1], k - s.org[2]};
    const int along = s.axis == 0 ? i : (s.axis == 1 ? j : k);
    p = s.minus ? s.pend - along : along - s.pend;
    c[s.axis] = p;
    i1 = ((uint32_t)c[0] * s.d1[1] + c[1]) * s.d1[2] + c[2];
    i2 = ((uint32_t)c[0] * s.d2[1] + c[1]) * s.d2[2] + c[2];
    o1 = (uint32_t)s.d1[0] * s.d1[1] * s.d1[2];
    o2 = (uint32_t)s.d2[0] * s.d2[1] * s.d2[2];
}

// Phi of the first slab in `smask` for cell (i, j, k), loaded as soon as the
// cell's slabs are known: the loads then run beside the cell's field loads
// instead of after the main update that the correction follows.
struct PhiPre {
    int s;             // the slab, -1 when none
    int p;
    uint32_t q1, q2, o1, o2;
    float a1, b1, a2, b2;  // Phi1 orders 0 and 1, Phi2 orders 0 and 1
};

__device__ __forceinline__ void phi_preload(const SlabArgs *slab, uint32_t smask, int i, int j, int k, PhiPre &pp)
{
    pp.s = -1;
    pp.a1 = pp.b1 = pp.a2 = pp.b2 = 0.0f;
    if (!smask) return;
    const int s = __ffs(smask) - 1;
    const SlabArgs &sl = slab[s];
    slab_index(sl, i, j, k, pp.p, pp.q1, pp.q2, pp.o1, pp.o2);
    pp.s = s;
    const gfloat_t *f1 = (const gfloat_t *)sl.phi1, *f2 = (const gfloat_t *)sl.phi2;
    pp.a1 = f1[pp.q1];
    pp.a2 = f2[pp.q2];
    if (sl.order == 2) {
        pp.b1 = f1[pp.q1 + pp.o1];
        pp.b2 = f2[pp.q2 + pp.o2];
    }
}

// Profile index, Phi indices and Phi values of slab `s` for cell (i, j, k):
// from the preload when it holds that slab, else loaded here.
__device__ __forceinline__ void phi_get(const SlabArgs &sl, int s, const PhiPre &pre, int i, int j, int k, int &p,
                                        uint32_t &q1, uint32_t &q2, uint32_t &o1, uint32_t &o2, float &a1,
                                        float &b1, float &a2, float &b2)
{
    if (s == pre.s) {
        p = pre.p;
        q1 = pre.q1; q2 = pre.q2; o1 = pre.o1; o2 = pre.o2;
        a1 = pre.a1; b1 = pre.b1; a2 = pre.a2; b2 = pre.b2;
        return;
    }
    slab_index(sl, i, j, k, p, q1, q2, o1, o2);
    const gfloat_t *f1 = (const gfloat_t *)sl.phi1, *f2 = (const gfloat_t *)sl.phi2;
    a1 = f1[q1];
    a2 = f2[q2];
    b1 = sl.order == 2 ? f1[q1 + o1] : 0.0f;
    b2 = sl.order == 2 ? f2[q2 + o2] : 0.0f;
}

// Persistent grid: blocks are dealt to the 8 XCDs round robin, so block b is
// renumbered (b mod 8) * per_xcd + b / 8; in each round an XCD then works on a
// contiguous run of items, i.e. neighbouring plane segments that share rows
// through its L2.
__device__ __forceinline__ int logical_block(int per_xcd)
{
    const int b = blockIdx.x;
    return (b & 7) * per_xcd + (b >> 3);
}

// Copy a small table to LDS with four loads in flight per thread.
// Slab descriptions go to LDS so the PML branch reads them into VGPRs when a
// lane needs them instead of holding them in SGPRs for the whole loop.
__device__ __forceinline__ void stage_slabs(SlabArgs *dst, const SlabArgs *slab, int nslab)
{
    const int words = (int)(sizeof(SlabArgs) / 4) * nslab;
    const int *src = reinterpret_cast<const int *>(slab);
    int *d = reinterpret_cast<int *>(dst);
    for (int q = threadIdx.x; q < words; q += blockDim.x) d[q] = src[q];
}

__device__ __forceinline__ void stage(float *dst, const float *src, int n)
{
    const int B = blockDim.x;
    for (int q = threadIdx.x; q < n; q += 4 * B) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = q + u * B < n ? src[q + u * B] : 0.f;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (q + u * B < n) dst[q + u * B] = v[u];
    }
}

// Coefficients in LDS: (CA, CBx, CBy, CBz) or (DA, DBx, DBy, DBz) as one
// float4 per material, then srce/srcm as one float per material.
__device__ __forceinline__ void stage_coef(float4 *c4, float *c1, const float *src, int nmat)
{
    for (int m = threadIdx.x; m < nmat; m += blockDim.x) {
        const float *r = src + 5 * m;
        c4[m] = make_float4(r[0], r[1], r[2], r[3]);
        c1[m] = r[4];
    }
}

// Dispersive part A for one component (gprMax/fields_updates_ext.pyx:113-296):
// returns phi and advances the P auxiliary values from the old field e.
template <int P, bool CPLX>
__device__ __forceinline__ float disp_a(const float *D, uint32_t m, float *tr, float *ti, float e)
{
    float phi = 0.f;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (CPLX) {
            const float *c = D + ((size_t)m * 3 * P + 3 * p) * 2;
            if (P == 1) phi = c[0] * tr[p];
            else phi = phi + c[0] * tr[p];
            const float er = c[2], ei = c[3], zr = c[4], zi = c[5];
            const float nr = (er * tr[p] - ei * ti[p]) + zr * e;
            const float ni = (er * ti[p] + ei * tr[p]) + zi * e;
            tr[p] = nr;
            ti[p] = ni;
        } else {
            const float *c = D + (size_t)m * 3 * P + 3 * p;
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
            const float *c = D + ((size_t)m * 3 * P + 3 * p) * 2;
            tr[p] = tr[p] - c[4] * e;
            ti[p] = ti[p] - c[5] * e;
        } else {
            tr[p] = tr[p] - D[(size_t)m * 3 * P + 3 * p + 2] * e;
        }
    }
}

// Dispersive auxiliary values, compacted: component c keeps its P poles
// only for the edges whose material has poles, in cell order,
// T[c][p * n[c] + slot], with slot[c][cell] the edge's place in that order
// (read only on such edges).
struct TArgs {
    float *T[3];
    const uint32_t *slot[3];
    long long n[3];
};

// Slot of an edge whose material has no poles.
constexpr uint32_t T_NOSLOT = 0xffffffffu;

// Loads the P values of component c in slot s.
template <int P, bool CPLX>
__device__ __forceinline__ void t_load_at(const TArgs &t, int c, uint32_t s, float *tr, float *ti)
{
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const size_t e = (size_t)p * (size_t)t.n[c] + s;
        if (CPLX) {
            const float2 v = ld(reinterpret_cast<const float2 *>(t.T[c]) + e, 0u);
            tr[p] = v.x;
            ti[p] = v.y;
        } else {
            tr[p] = ld((const float *)t.T[c] + e, 0u);
            ti[p] = 0.f;
        }
    }
}

// Loads the P values of edge (c, po + to / 4) and returns its slot.
template <int P, bool CPLX>
__device__ __forceinline__ uint32_t t_load(const TArgs &t, int c, size_t po, uint32_t to, float *tr, float *ti)
{
    const uint32_t s = ld(t.slot[c] + po, to);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const size_t e = (size_t)p * (size_t)t.n[c] + s;
        if (CPLX) {
            const float2 v = ld(reinterpret_cast<const float2 *>(t.T[c]) + e, 0u);
            tr[p] = v.x;
            ti[p] = v.y;
        } else {
            tr[p] = ld((const float *)t.T[c] + e, 0u);
            ti[p] = 0.f;
        }
    }
    return s;
}

template <int P, bool CPLX>
__device__ __forceinline__ void t_store(const TArgs &t, int c, uint32_t s, const float *tr, const float *ti)
{
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const size_t e = (size_t)p * (size_t)t.n[c] + s;
        if (CPLX)
            st(reinterpret_cast<float2 *>(t.T[c]) + e, 0u, make_float2(tr[p], ti[p]));
        else
            st(t.T[c] + e, 0u, tr[p]);
    }
}

// ---------------------------------------------------------------- fused step
// One whole iteration -- the H pass and then the E pass -- in one kernel.  The
// state is double-buffered: fields and H-PML Phi are read from the current
// arrays and written to the next ones, so no block ever reads what another
// block writes in the same launch and blocks need no ordering.  (Dispersive T
// and E-PML Phi are read and written only by the owner of the cell, in place.)
//
// Work item = (tile, x chunk).  A tile is a run of L consecutive cells of the
// flattened (j, k) plane, preceded by the nz+1 cells before it: that halo holds
// the (j-1) and (k-1) neighbours of every tile cell.  A block of B threads
// holds B*CPT = nz+1+L cells, thread t the cells t, t+B, ...  For each plane i
// of the chunk, and first for the plane before it, the block computes new H
// for all its cells (the halo cells and the halo plane are recomputed, never
// written), shares them through LDS, and computes new E for the tile cells.
// The new Hy, Hz of plane i-1 that E at plane i needs stay in registers.
// Sources and receivers of the fused step.  They are read only on planes that
// hold one, from device memory, so their addresses do not occupy registers
// for the whole kernel.
struct StepRare {
    SrcArgs srcM, srcV, srcD;    // magnetic dipoles, voltage sources, Hertzian dipoles
    int nrx;
    const int *rxc;              // [nrx][3]
    float *rxbuf;                // [6][iterations][nrx]
};

struct FusedArgs {
    int nx, ny, nz, plane;
    int L, ntiles, nitems, per_xcd, xc;
    const float *Fc;             // current state: Ex, Ey, Ez, Hx, Hy, Hz, `cs` bytes apart
    float *Fx;                   // next state, same layout
    uint32_t cs;                 // bytes between the components of a state
    const void *ID;              // [6][ntot] compact material indices
    long long ntot;
    const float *coefE, *coefH;
    const float *RE, *RH;        // all slabs' R tables, E and H
    int nR;
    const float *disp;
    const uint8_t *dispflag;
    TArgs T;
    int nmat, form, nslab;
    SlabArgs slabE[MAX_SLABS], slabH[MAX_SLABS];  // slabH: Phi read current, written next
    int it, iterations;
    const StepRare *rare;
    const uint32_t *planeflag;
};

// Pointers into the dynamic LDS of the fused step: both coefficient tables,
// both R tables, the dispersive coefficients, the H exchange area of the
// block's cells and the dispersive flags; the slab descriptions are static.
struct StepLds {
    float4 *cE, *cH;
    float *sE, *sH, *lRE, *lRH, *ldisp, *lHx, *lHy, *lHz;
    uint8_t *lflag;
    SlabArgs *lsE, *lsH;
};

template <int P, bool CPLX, int CPT>
__device__ __forceinline__ StepLds step_layout(float4 *lds4, SlabArgs *lsE, SlabArgs *lsH, const FusedArgs &a)
{
    StepLds S;
    const int nm = a.nmat, TC = blockDim.x * CPT;
    const int nd = P > 0 ? nm * 3 * P * (CPLX ? 2 : 1) : 0;
    S.cE = lds4;
    S.cH = lds4 + nm;
    S.sE = reinterpret_cast<float *>(lds4 + 2 * nm);
    S.sH = S.sE + nm;
    S.lRE = S.sH + nm;
    S.lRH = S.lRE + a.nR;
    S.ldisp = S.lRH + a.nR;
    S.lHx = S.ldisp + nd;
    S.lHy = S.lHx + TC;
    S.lHz = S.lHy + TC;
    S.lflag = reinterpret_cast<uint8_t *>(S.lHz + TC);
    S.lsE = lsE;
    S.lsH = lsH;
    return S;
}

// Copies one model's tables to LDS (the caller synchronises).
template <int P, bool CPLX>
__device__ __forceinline__ void step_stage(const StepLds &S, const FusedArgs &a)
{
    const int nm = a.nmat;
    stage_coef(S.cE, S.sE, a.coefE, nm);
    stage_coef(S.cH, S.sH, a.coefH, nm);
    stage(S.lRE, a.RE, a.nR);
    stage(S.lRH, a.RH, a.nR);
    stage_slabs(S.lsE, a.slabE, a.nslab);
    stage_slabs(S.lsH, a.slabH, a.nslab);
    if (P > 0) {
        stage(S.ldisp, a.disp, nm * 3 * P * (CPLX ? 2 : 1));
        for (int q = threadIdx.x; q < nm; q += blockDim.x) S.lflag[q] = a.dispflag[q];
    }
}

// One work item (tile, x chunk) of iteration `it` of one model.
template <class IDT, int P, bool CPLX, int FORM, int ORDER, int CPT>
__device__ __forceinline__ void step_item(const FusedArgs &a, const StepLds &S, int item, int it)
{
    const int B = blockDim.x;
    const float4 *cE = S.cE, *cH = S.cH;
    const float *sE = S.sE, *sH = S.sH, *lRE = S.lRE, *lRH = S.lRH, *ldisp = S.ldisp;
    float *lHx = S.lHx, *lHy = S.lHy, *lHz = S.lHz;
    const uint8_t *lflag = S.lflag;
    const SlabArgs *lsE = S.lsE, *lsH = S.lsH;
    (void)ldisp;
    (void)lflag;

    const int nx = a.nx, ny = a.ny, nz = a.nz, NZ1 = nz + 1;
    const size_t N = a.ntot, Pl = a.plane;
    // Three buffer resources cover everything streamed: the current state,
    // the next state and the material indices; the component is part of the
    // scalar offset.
    const rsrc_t rF = rsrc(a.Fc), wF = rsrc(a.Fx), rI = rsrc(a.ID);
    const uint32_t cs = a.cs, ci = (uint32_t)N * (uint32_t)sizeof(IDT);
    const uint32_t P4 = (uint32_t)Pl * 4u, Pid = (uint32_t)Pl * (uint32_t)sizeof(IDT);
    {
        const int tile = item % a.ntiles, chunk = item / a.ntiles;
        const int q0 = tile * a.L - NZ1;
        const int qe = min(tile * a.L + a.L, a.plane);
        const int xa = chunk * a.xc, xb = min(xa + a.xc, nx + 1);
        const int xs = xa > 0 ? xa - 1 : xa;

        uint32_t to[CPT];      // byte offset of the cell within a plane (0 for dead lanes)
        uint32_t jk[CPT];      // j << 16 | k
        // bit 0: live cell; bit 1: tile cell (not halo); bits 2..7: slabs whose
        // (j, k) range holds the cell for E, bits 8..13 the same for H
        uint32_t st[CPT];
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int t = threadIdx.x + u * B;
            const int q = q0 + t;
            const bool live = q >= 0 && q < qe;
            const int qq = live ? q : 0;
            const int j = qq / NZ1, k = qq - j * NZ1;
            to[u] = (uint32_t)qq * 4u;
            jk[u] = ((uint32_t)j << 16) | (uint32_t)k;
            uint32_t me = 0, mh = 0;
            for (int s = 0; ORDER > 0 && s < a.nslab; ++s) {
                if (j >= lsE[s].lo[1] && j < lsE[s].hi[1] && k >= lsE[s].lo[2] && k < lsE[s].hi[2]) me |= 1u << s;
                if (j >= lsH[s].lo[1] && j < lsH[s].hi[1] && k >= lsH[s].lo[2] && k < lsH[s].hi[2]) mh |= 1u << s;
            }
            st[u] = live ? (1u | (t >= NZ1 ? 2u : 0u) | (me << 2) | (mh << 8)) : 0u;
        }

        // The loads of one plane for one cell: old E of the next plane, old H
        // and material indices of this plane, E neighbours in y and z.
        struct PlaneLoad {
            float ex1, ey1, ez1, hx, hy, hz, ezj, exj, eyk, exk;
            uint32_t m0, m1, m2, m3, m4, m5;
        };
        auto load = [&](int ip, PlaneLoad *L) {
            const uint32_t so = (uint32_t)ip * P4;
            const uint32_t sn = (uint32_t)(ip < nx ? ip + 1 : ip) * P4;
            const uint32_t si = (uint32_t)ip * Pid;
#pragma unroll
            for (int u = 0; u < CPT; ++u) {
                const int j = (int)(jk[u] >> 16), k = (int)(jk[u] & 0xffffu);
                // Neighbours past the grid edge are clamped to the cell itself
                // (a +0 difference, which no update applied there uses).
                const uint32_t toj = j < ny ? to[u] + 4u * NZ1 : to[u];
                const uint32_t tok = k < nz ? to[u] + 4u : to[u];
                const uint32_t ti = (to[u] >> 2) * (uint32_t)sizeof(IDT);
                L[u].ex1 = bld(rF, to[u], sn);
                L[u].ey1 = bld(rF, to[u], sn + cs);
                L[u].ez1 = bld(rF, to[u], sn + 2 * cs);
                L[u].hx = bld(rF, to[u], so + 3 * cs);
                L[u].hy = bld(rF, to[u], so + 4 * cs);
                L[u].hz = bld(rF, to[u], so + 5 * cs);
                L[u].ezj = bld(rF, toj, so + 2 * cs);
                L[u].exj = bld(rF, toj, so);
                L[u].eyk = bld(rF, tok, so + cs);
                L[u].exk = bld(rF, tok, so);
                L[u].m0 = bld_id<IDT>(rI, ti, si);
                L[u].m1 = bld_id<IDT>(rI, ti, si + ci);
                L[u].m2 = bld_id<IDT>(rI, ti, si + 2 * ci);
                L[u].m3 = bld_id<IDT>(rI, ti, si + 3 * ci);
                L[u].m4 = bld_id<IDT>(rI, ti, si + 4 * ci);
                L[u].m5 = bld_id<IDT>(rI, ti, si + 5 * ci);
            }
        };

        float ex[CPT], ey[CPT], ez[CPT], hym[CPT], hzm[CPT];
        {
            const uint32_t so = (uint32_t)xs * P4;
#pragma unroll
            for (int u = 0; u < CPT; ++u) {
                ex[u] = bld(rF, to[u], so);
                ey[u] = bld(rF, to[u], so + cs);
                ez[u] = bld(rF, to[u], so + 2 * cs);
                hym[u] = 0.f;
                hzm[u] = 0.f;
            }
        }
        PlaneLoad cur[CPT];
        load(xs, cur);

        for (int i = xs; i < xb; ++i) {
            const bool hp = i < xa;  // halo plane: H only, nothing written
            const uint32_t pf = ld(a.planeflag, 4u * (uint32_t)i);
            const uint32_t so = (uint32_t)i * P4;

            float ex1[CPT], ey1[CPT], ez1[CPT], hx[CPT], hy[CPT], hz[CPT];
            float ezj[CPT], exj[CPT], eyk[CPT], exk[CPT];
            uint32_t m0[CPT], m1[CPT], m2[CPT], m3[CPT], m4[CPT], m5[CPT];
#pragma unroll
            for (int u = 0; u < CPT; ++u) {
                ex1[u] = cur[u].ex1; ey1[u] = cur[u].ey1; ez1[u] = cur[u].ez1;
                hx[u] = cur[u].hx; hy[u] = cur[u].hy; hz[u] = cur[u].hz;
                ezj[u] = cur[u].ezj; exj[u] = cur[u].exj; eyk[u] = cur[u].eyk; exk[u] = cur[u].exk;
                m0[u] = cur[u].m0; m1[u] = cur[u].m1; m2[u] = cur[u].m2;
                m3[u] = cur[u].m3; m4[u] = cur[u].m4; m5[u] = cur[u].m5;
            }
            // The next plane's loads fly during this whole plane: both updates
            // and both barriers.
            if (i + 1 < xb) load(i + 1, cur);

            // ---- H: update_magnetic, H-PML slabs in order, magnetic dipoles
#pragma unroll
            for (int u = 0; u < CPT; ++u) {
                const int j = (int)(jk[u] >> 16), k = (int)(jk[u] & 0xffffu);
                const bool own = (st[u] & 2u) && !hp;
                const float dEz_y = ezj[u] - ez[u], dEy_z = eyk[u] - ey[u], dEx_z = exk[u] - ex[u];
                const float dEz_x = ez1[u] - ez[u], dEy_x = ey1[u] - ey[u], dEx_y = exj[u] - ex[u];
                float nhx = hx[u], nhy = hy[u], nhz = hz[u];
                if (i >= 1 && j < ny && k < nz) {
                    const float4 q = cH[m3[u]];
                    nhx = q.x * nhx - q.z * dEz_y + q.w * dEy_z;
                }
                if (i < nx && j >= 1 && k < nz) {
                    const float4 q = cH[m4[u]];
                    nhy = q.x * nhy - q.w * dEx_z + q.y * dEz_x;
                }
                if (i < nx && j < ny && k >= 1) {
                    const float4 q = cH[m5[u]];
                    nhz = q.x * nhz - q.y * dEy_x + q.z * dEx_y;
                }
                const uint32_t smask = (st[u] >> 8) & (pf >> PF_SLAB_H) & 63u;
                if (ORDER > 0 && smask) {
                    for (int s = 0; s < a.nslab; ++s) {
                        if (!(smask & (1u << s))) continue;
                        const SlabArgs &sl = lsH[s];
                        int p;
                        uint32_t q1, q2, o1, o2;
                        slab_index(sl, i, j, k, p, q1, q2, o1, o2);
                        const float *R = lRH + sl.roff;
                        if (sl.axis == 0) {
                            nhy = pml_t<FORM, ORDER>(sl, R, p, nhy, true, sH[m4[u]], dEz_x / sl.d, sl.phi1, sl.phw1, own, q1, o1);
                            nhz = pml_t<FORM, ORDER>(sl, R, p, nhz, false, sH[m5[u]], dEy_x / sl.d, sl.phi2, sl.phw2, own, q2, o2);
                        } else if (sl.axis == 1) {
                            nhx = pml_t<FORM, ORDER>(sl, R, p, nhx, false, sH[m3[u]], dEz_y / sl.d, sl.phi1, sl.phw1, own, q1, o1);
                            nhz = pml_t<FORM, ORDER>(sl, R, p, nhz, true, sH[m5[u]], dEx_y / sl.d, sl.phi2, sl.phw2, own, q2, o2);
                        } else {
                            nhx = pml_t<FORM, ORDER>(sl, R, p, nhx, true, sH[m3[u]], dEy_z / sl.d, sl.phi1, sl.phw1, own, q1, o1);
                            nhy = pml_t<FORM, ORDER>(sl, R, p, nhy, false, sH[m4[u]], dEx_z / sl.d, sl.phi2, sl.phw2, own, q2, o2);
                        }
                    }
                }
                if ((pf & PF_SRCH) && (st[u] & 1u)) {  // magnetic dipoles, gprMax/sources.py:202-232
                    const SrcArgs &M = a.rare->srcM;
                    for (int s = 0; s < M.n; ++s) {
                        const int *in = M.info + 6 * s;
                        if (in[0] != i || it < in[4] || it > in[5]) continue;
                        if (in[1] == j && in[2] == k) {
                            const int pol = in[3];
                            const uint32_t m = pol == 0 ? m3[u] : (pol == 1 ? m4[u] : m5[u]);
                            const float w = M.waves[(size_t)s * a.iterations + it];
                            const float v = sH[m] * w * M.fac[2 * s];
                            if (pol == 0) nhx = nhx - v;
                            else if (pol == 1) nhy = nhy - v;
                            else nhz = nhz - v;
                        }
                    }
                }
                if (own && (pf & PF_RX)) {  // the state at the start of the iteration
                    const StepRare &Q = *a.rare;
                    for (int r = 0; r < Q.nrx; ++r) {
                        if (Q.rxc[3 * r] != i || Q.rxc[3 * r + 1] != j || Q.rxc[3 * r + 2] != k) continue;
                        float *o = Q.rxbuf + (size_t)it * Q.nrx + r;
                        const size_t cs = (size_t)a.iterations * Q.nrx;
                        o[0] = ex[u]; o[cs] = ey[u]; o[2 * cs] = ez[u];
                        o[3 * cs] = hx[u]; o[4 * cs] = hy[u]; o[5 * cs] = hz[u];
                    }
                }
                if (own) {
                    bst(wF, to[u], so + 3 * cs, nhx);
                    bst(wF, to[u], so + 4 * cs, nhy);
                    bst(wF, to[u], so + 5 * cs, nhz);
                }
                const int t = threadIdx.x + u * B;
                lHx[t] = nhx;
                lHy[t] = nhy;
                lHz[t] = nhz;
                hx[u] = nhx;
                hy[u] = nhy;
                hz[u] = nhz;
            }
            __syncthreads();

            // ---- E: update_electric or dispersive A, E-PML slabs in order,
            // voltage sources, Hertzian dipoles, dispersive B
            if (!hp) {
#pragma unroll
                for (int u = 0; u < CPT; ++u) {
                    if (!(st[u] & 2u)) continue;
                    const int j = (int)(jk[u] >> 16), k = (int)(jk[u] & 0xffffu);
                    const int t = threadIdx.x + u * B;
                    const float hzj = lHz[t - NZ1], hxj = lHx[t - NZ1];
                    const float hyk = lHy[t - 1], hxk = lHx[t - 1];
                    const float dHz_y = hz[u] - hzj, dHy_z = hy[u] - hyk, dHx_z = hx[u] - hxk;
                    const float dHz_x = hz[u] - hzm[u], dHy_x = hy[u] - hym[u], dHx_y = hx[u] - hxj;
                    float nex = ex[u], ney = ey[u], nez = ez[u];
                    const size_t po = (size_t)i * Pl;
                    float txr[P > 0 ? P : 1], txi[P > 0 ? P : 1];
                    float tyr[P > 0 ? P : 1], tyi[P > 0 ? P : 1];
                    float tzr[P > 0 ? P : 1], tzi[P > 0 ? P : 1];
                    uint32_t sx = 0, sy = 0, sz = 0;
                    const bool rx_ = i < nx && j >= 1 && j < ny && k >= 1 && k < nz;
                    const bool ry_ = i >= 1 && i < nx && j < ny && k >= 1 && k < nz;
                    const bool rz_ = i >= 1 && i < nx && j >= 1 && j < ny && k < nz;
                    const bool dx_ = P > 0 && rx_ && lflag[m0[u]];
                    const bool dy_ = P > 0 && ry_ && lflag[m1[u]];
                    const bool dz_ = P > 0 && rz_ && lflag[m2[u]];
                    if (P > 0) {
                        if (dx_) sx = t_load<P, CPLX>(a.T, 0, po, to[u], txr, txi);
                        if (dy_) sy = t_load<P, CPLX>(a.T, 1, po, to[u], tyr, tyi);
                        if (dz_) sz = t_load<P, CPLX>(a.T, 2, po, to[u], tzr, tzi);
                    }
                    if (rx_) {
                        const float4 q = cE[m0[u]];
                        float v = q.x * nex + q.z * dHz_y - q.w * dHy_z;
                        if (P > 0) {
                            const float phi = dx_ ? disp_a<P, CPLX>(ldisp, m0[u], txr, txi, nex) : 0.f;
                            v = v - sE[m0[u]] * phi;
                        }
                        nex = v;
                    }
                    if (ry_) {
                        const float4 q = cE[m1[u]];
                        float v = q.x * ney + q.w * dHx_z - q.y * dHz_x;
                        if (P > 0) {
                            const float phi = dy_ ? disp_a<P, CPLX>(ldisp, m1[u], tyr, tyi, ney) : 0.f;
                            v = v - sE[m1[u]] * phi;
                        }
                        ney = v;
                    }
                    if (rz_) {
                        const float4 q = cE[m2[u]];
                        float v = q.x * nez + q.y * dHy_x - q.z * dHx_y;
                        if (P > 0) {
                            const float phi = dz_ ? disp_a<P, CPLX>(ldisp, m2[u], tzr, tzi, nez) : 0.f;
                            v = v - sE[m2[u]] * phi;
                        }
                        nez = v;
                    }
                    const uint32_t smask = (st[u] >> 2) & (pf >> PF_SLAB_E) & 63u;
                    if (ORDER > 0 && smask) {
                        for (int s = 0; s < a.nslab; ++s) {
                            if (!(smask & (1u << s))) continue;
                            const SlabArgs &sl = lsE[s];
                            int p;
                            uint32_t q1, q2, o1, o2;
                            slab_index(sl, i, j, k, p, q1, q2, o1, o2);
                            const float *R = lRE + sl.roff;
                            if (sl.axis == 0) {
                                ney = pml_t<FORM, ORDER>(sl, R, p, ney, false, sE[m1[u]], dHz_x / sl.d, sl.phi1, sl.phi1, true, q1, o1);
                                nez = pml_t<FORM, ORDER>(sl, R, p, nez, true, sE[m2[u]], dHy_x / sl.d, sl.phi2, sl.phi2, true, q2, o2);
                            } else if (sl.axis == 1) {
                                nex = pml_t<FORM, ORDER>(sl, R, p, nex, true, sE[m0[u]], dHz_y / sl.d, sl.phi1, sl.phi1, true, q1, o1);
                                nez = pml_t<FORM, ORDER>(sl, R, p, nez, false, sE[m2[u]], dHx_y / sl.d, sl.phi2, sl.phi2, true, q2, o2);
                            } else {
                                nex = pml_t<FORM, ORDER>(sl, R, p, nex, false, sE[m0[u]], dHy_z / sl.d, sl.phi1, sl.phi1, true, q1, o1);
                                ney = pml_t<FORM, ORDER>(sl, R, p, ney, true, sE[m1[u]], dHx_z / sl.d, sl.phi2, sl.phi2, true, q2, o2);
                            }
                        }
                    }
                    if (pf & PF_SRCE) {
                        // Voltage sources, then Hertzian dipoles (gprMax/model_build_run.py:458-461).
                        const SrcArgs &V = a.rare->srcV, &D = a.rare->srcD;
                        for (int s = 0; s < V.n; ++s) {  // gprMax/sources.py:81-120
                            const int *in = V.info + 6 * s;
                            if (in[0] != i || it < in[4] || it > in[5]) continue;
                            if (in[1] == j && in[2] == k) {
                                const int pol = in[3] & 3;
                                const bool hard = in[3] >= 4;
                                const uint32_t m = pol == 0 ? m0[u] : (pol == 1 ? m1[u] : m2[u]);
                                const float w = V.waves[(size_t)s * a.iterations + it];
                                float *F = pol == 0 ? &nex : (pol == 1 ? &ney : &nez);
                                if (hard) *F = -w / V.fac[2 * s];
                                else *F = *F - sE[m] * w * V.fac[2 * s];
                            }
                        }
                        for (int s = 0; s < D.n; ++s) {  // gprMax/sources.py:163-193
                            const int *in = D.info + 6 * s;
                            if (in[0] != i || it < in[4] || it > in[5]) continue;
                            if (in[1] == j && in[2] == k) {
                                const int pol = in[3];
                                const uint32_t m = pol == 0 ? m0[u] : (pol == 1 ? m1[u] : m2[u]);
                                const float w = D.waves[(size_t)s * a.iterations + it];
                                const float v = sE[m] * w * D.fac[2 * s] * D.fac[2 * s + 1];
                                if (pol == 0) nex = nex - v;
                                else if (pol == 1) ney = ney - v;
                                else nez = nez - v;
                            }
                        }
                    }
                    if (P > 0) {
                        if (dx_) { disp_b<P, CPLX>(ldisp, m0[u], txr, txi, nex); t_store<P, CPLX>(a.T, 0, sx, txr, txi); }
                        if (dy_) { disp_b<P, CPLX>(ldisp, m1[u], tyr, tyi, ney); t_store<P, CPLX>(a.T, 1, sy, tyr, tyi); }
                        if (dz_) { disp_b<P, CPLX>(ldisp, m2[u], tzr, tzi, nez); t_store<P, CPLX>(a.T, 2, sz, tzr, tzi); }
                    }
                    bst(wF, to[u], so, nex);
                    bst(wF, to[u], so + cs, ney);
                    bst(wF, to[u], so + 2 * cs, nez);
                }
            }
#pragma unroll
            for (int u = 0; u < CPT; ++u) {
                hym[u] = hy[u];
                hzm[u] = hz[u];
                ex[u] = ex1[u];
                ey[u] = ey1[u];
                ez[u] = ez1[u];
            }
            __syncthreads();
        }
    }
}

template <class IDT, int P, bool CPLX, int FORM, int ORDER, int CPT>
__global__ void __launch_bounds__(1024) k_step(FusedArgs a)
{
    extern __shared__ float4 lds4[];
    __shared__ SlabArgs lsE[MAX_SLABS], lsH[MAX_SLABS];
    const StepLds S = step_layout<P, CPLX, CPT>(lds4, lsE, lsH, a);
    step_stage<P, CPLX>(S, a);
    __syncthreads();
    for (int item = logical_block(a.per_xcd); item < a.nitems; item += gridDim.x)
        step_item<IDT, P, CPLX, FORM, ORDER, CPT>(a, S, item, a.it);
}

// The fused step of many models in one launch.  `args` holds every model's
// arguments twice, [2][ntr]: for its current state and with the two states
// exchanged; `flip` picks the row, so consecutive iterations alternate rows
// without host updates.  Model t owns items [item0[t], item0[t+1]).
struct BatchArgs {
    const FusedArgs *args;
    const int *item0;
    int ntr, nitems, per_xcd;
    int flip, it;
};

template <class IDT, int P, bool CPLX, int FORM, int ORDER, int CPT>
__global__ void __launch_bounds__(1024) k_step_batch(BatchArgs b)
{
    extern __shared__ float4 lds4[];
    __shared__ SlabArgs lsE[MAX_SLABS], lsH[MAX_SLABS];
    int staged = -1;
    for (int g = logical_block(b.per_xcd); g < b.nitems; g += gridDim.x) {
        int lo = 0, hi = b.ntr - 1;  // last model whose first item <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (b.item0[mid] <= g) lo = mid;
            else hi = mid - 1;
        }
        const FusedArgs &a = b.args[b.flip * b.ntr + lo];
        const StepLds S = step_layout<P, CPLX, CPT>(lds4, lsE, lsH, a);
        if (lo != staged) {
            __syncthreads();
            step_stage<P, CPLX>(S, a);
            __syncthreads();
            staged = lo;
        }
        step_item<IDT, P, CPLX, FORM, ORDER, CPT>(a, S, g - b.item0[lo], b.it);
    }
}

typedef void (*step_kernel)(FusedArgs);
typedef void (*batch_kernel)(BatchArgs);

// Cells per thread of the fused step: a block of B threads holds B x this.
constexpr int STEP_CELLS_PER_THREAD = 1;

// Fused kernels (one model per launch, and many) for the given
// material-index width, poles and coefficient kind.
struct StepKernels {
    step_kernel one;
    batch_kernel many;
};

template <int FORM, int ORDER>
StepKernels step_kernel_table(int id_bytes, int P, bool C)
{
    constexpr int K = STEP_CELLS_PER_THREAD;
#define GPX_SK(IDT, PP, CC) \
    StepKernels{k_step<IDT, PP, CC, FORM, ORDER, K>, k_step_batch<IDT, PP, CC, FORM, ORDER, K>}
    (void)id_bytes;  // material indices are always uint16 on the device
    switch (P) {
    case 0: return GPX_SK(uint16_t, 0, false);
    case 1: return C ? GPX_SK(uint16_t, 1, true) : GPX_SK(uint16_t, 1, false);
    case 2: return C ? GPX_SK(uint16_t, 2, true) : GPX_SK(uint16_t, 2, false);
    case 3: return C ? GPX_SK(uint16_t, 3, true) : GPX_SK(uint16_t, 3, false);
    default: return C ? GPX_SK(uint16_t, 4, true) : GPX_SK(uint16_t, 4, false);
    }
#undef GPX_SK
}

}  // namespace gpx

#endif  // GPX_STEP_H
