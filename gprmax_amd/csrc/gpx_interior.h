// Fused interior kernel of the split step (gpx_set_step form 3): one sweep
// that computes new H and then new E for the cells of an interior box, from
// the current state into the next one (double-buffered), for models without
// dispersive materials.
//
// The E update of cell (i, j, k) needs new H at (i, j, k) and at its i-1, j-1
// and k-1 neighbours.  A lane owns one (j, k) column of the box and marches x
// through a chunk of planes, so new H at i-1 is its own from the previous
// plane (the chunk starts one plane early to make it).  A block holds
// INT_BLOCK consecutive cells of the box's (j, k) rectangle; each plane's new
// H goes to LDS (two buffers, planes alternating, one barrier per plane), and
// new H at j-1 and k-1 is read there from the lanes Kin and 1 back.  Where
// that lane is not in the block (the block's first box row for j-1; its first
// lane and the first cell of a box row for k-1) the lane recomputes the value
// from the current state: the same expression on the same inputs gives the
// same bits as the owning lane's result.  The box leaves one plain cell on its
// minus sides, so every recomputed neighbour is a cell of the plain update (no
// PML, all neighbours inside the grid) and needs no term this kernel lacks.
//
// The formulas and their association are those of update_magnetic and
// update_electric (gprMax/fields_updates_ext.pyx:81-107, 404-412), then
// magnetic dipoles on H (also on recomputed neighbours) and voltage sources
// then Hertzian dipoles on E (gprMax/model_build_run.py:444-461).  The
// receivers sample the current state.
#pragma once

namespace gpx {

constexpr int INT_BLOCK = 512;  // 256: 281 us, 512: 249 us, 1024: 254 us on the 300x300x200 half-space

struct IntArgs {
    int nx, ny, nz, plane;
    int box[6];                // [i0, i1) x [j0, j1) x [k0, k1)
    int nseg, nitems, per_xcd, xc;
    const float *Fc;           // current state: component c at byte offset c * cs
    float *Fx;                 // next state
    uint32_t cs;
    const void *ID;            // [6][ntot]
    long long ntot;
    const float *coefE, *coefH;
    int nmat;
    int it, iterations;
    SrcArgs srcM, srcV, srcD;  // magnetic dipoles, voltage sources, Hertzian dipoles
    int nrx;
    const int *rxc;
    float *rxbuf;
    const uint32_t *planeflag;
};

// Magnetic dipoles of polarisation `pol` at cell (i, j, k) (gprMax/sources.py:202-232).
__device__ __forceinline__ float int_msrc(const IntArgs &a, const float *sH, int i, int j, int k, int pol,
                                          uint32_t m, float h, int it)
{
    for (int s = 0; s < a.srcM.n; ++s) {
        const int *in = a.srcM.info + 6 * s;
        if (in[0] != i || in[1] != j || in[2] != k || in[3] != pol || it < in[4] || it > in[5]) continue;
        const float w = a.srcM.waves[(size_t)s * a.iterations + it];
        h = h - sH[m] * w * a.srcM.fac[2 * s];
    }
    return h;
}

// The body for work item `item` (< a.nitems; block-uniform).
template <class IDT>
__device__ __forceinline__ void interior_run(const IntArgs &a, int item)
{
    extern __shared__ float4 lds4[];
    const int nm = a.nmat;
    float4 *cE = lds4, *cH = lds4 + nm;
    float *sE = reinterpret_cast<float *>(lds4 + 2 * nm), *sH = sE + nm;
    stage_coef(cE, sE, a.coefE, nm);
    stage_coef(cH, sH, a.coefH, nm);
    __syncthreads();

    const int it = a.it;
    const int B = blockDim.x;
    const int NZ1 = a.nz + 1;
    const int Kin = a.box[5] - a.box[4];
    const int ncell = (a.box[3] - a.box[2]) * Kin;
    const int seg = item % a.nseg, chunk = item / a.nseg;
    const int t = seg * B + (int)threadIdx.x;
    const bool live = t < ncell;
    const int q = live ? t : 0;
    const int j = a.box[2] + q / Kin, k = a.box[4] + q % Kin;
    const int xa = a.box[0] + chunk * a.xc, xb = min(xa + a.xc, a.box[1]);

    const rsrc_t rF = rsrc(a.Fc), wF = rsrc(a.Fx), rI = rsrc(a.ID);
    const uint32_t cs = a.cs, ci = (uint32_t)a.ntot * (uint32_t)sizeof(IDT);
    const uint32_t P4 = (uint32_t)a.plane * 4u, Pid = (uint32_t)a.plane * (uint32_t)sizeof(IDT);
    const uint32_t to = (uint32_t)(j * NZ1 + k) * 4u;
    const uint32_t tj = to + 4u * NZ1, tjm = to - 4u * NZ1, tk = to + 4u, tkm = to - 4u;
    const uint32_t tid = (to >> 2) * (uint32_t)sizeof(IDT), tidjm = tid - (uint32_t)(NZ1 * sizeof(IDT)),
                   tidkm = tid - (uint32_t)sizeof(IDT);
    enum { EX = 0, EY = 1, EZ = 2, HX = 3, HY = 4, HZ = 5 };
    auto C = [&](int c) { return (uint32_t)c * cs; };

    struct PL {
        float hx, hy, hz, ezj, exj, eyk, exk, ex1, ey1, ez1;
        uint32_t m0, m1, m2, m3, m4, m5;
    };
    __shared__ float lh[2][3][INT_BLOCK];  // new H of the block, planes alternating
    const int tl = threadIdx.x;
    const bool jrec = tl < Kin;
    // Loads of plane ip for the cell: its old H and E at i, i+1, j+1 and k+1,
    // and its six material indices.
    auto load = [&](int ip, PL &L) {
        const uint32_t so = (uint32_t)ip * P4, sn = so + P4, si = (uint32_t)ip * Pid;
        L.hx = bld(rF, to, so + C(HX));
        L.hy = bld(rF, to, so + C(HY));
        L.hz = bld(rF, to, so + C(HZ));
        L.ezj = bld(rF, tj, so + C(EZ));
        L.exj = bld(rF, tj, so + C(EX));
        L.eyk = bld(rF, tk, so + C(EY));
        L.exk = bld(rF, tk, so + C(EX));
        L.ex1 = bld(rF, to, sn + C(EX));
        L.ey1 = bld(rF, to, sn + C(EY));
        L.ez1 = bld(rF, to, sn + C(EZ));
        L.m0 = bld_id<IDT>(rI, tid, si);
        L.m1 = bld_id<IDT>(rI, tid, si + ci);
        L.m2 = bld_id<IDT>(rI, tid, si + 2 * ci);
        L.m3 = bld_id<IDT>(rI, tid, si + 3 * ci);
        L.m4 = bld_id<IDT>(rI, tid, si + 4 * ci);
        L.m5 = bld_id<IDT>(rI, tid, si + 5 * ci);
    };

    // the halo plane xa-1: new Hy and Hz there, for the E update of plane xa
    float ex, ey, ez, hym, hzm;
    {
        const int ih = xa - 1;
        const uint32_t so = (uint32_t)ih * P4, sn = so + P4, si = (uint32_t)ih * Pid;
        const float exh = bld(rF, to, so + C(EX)), eyh = bld(rF, to, so + C(EY)), ezh = bld(rF, to, so + C(EZ));
        const float hy0 = bld(rF, to, so + C(HY)), hz0 = bld(rF, to, so + C(HZ));
        const float exk = bld(rF, tk, so + C(EX)), exj = bld(rF, tj, so + C(EX));
        const float ey1 = bld(rF, to, sn + C(EY)), ez1 = bld(rF, to, sn + C(EZ));
        const uint32_t m4 = bld_id<IDT>(rI, tid, si + 4 * ci), m5 = bld_id<IDT>(rI, tid, si + 5 * ci);
        const uint32_t pf = ld(a.planeflag, 4u * (uint32_t)ih);
        float4 qy = cH[m4], qz = cH[m5];
        hym = qy.x * hy0 - qy.w * (exk - exh) + qy.y * (ez1 - ezh);
        hzm = qz.x * hz0 - qz.y * (ey1 - eyh) + qz.z * (exj - exh);
        if (pf & PF_SRCH) {
            hym = int_msrc(a, sH, ih, j, k, 1, m4, hym, it);
            hzm = int_msrc(a, sH, ih, j, k, 2, m5, hzm, it);
        }
        ex = bld(rF, to, sn + C(EX));
        ey = ey1;
        ez = ez1;
    }

    // No software prefetch: loading the next plane ahead raised the VGPR count
    // from under 64 to over 90 (5 waves per SIMD) for no gain.
    // Lanes that recompute a neighbour's new H load its inputs with the
    // plane, so the loads overlap the plane's own instead of following the
    // barrier.
    const bool krec = tl == 0 || k == a.box[4];
    PL cur;
    for (int i = xa; i < xb; ++i) {
        const uint32_t so = (uint32_t)i * P4;
        const uint32_t pf = ld(a.planeflag, 4u * (uint32_t)i);
        load(i, cur);
        float hxjm = 0.f, hzjm = 0.f, ezjm = 0.f, exjm = 0.f, eyjm = 0.f, eyjmk = 0.f, eyjm1 = 0.f;
        uint32_t mj3 = 0, mj5 = 0;
        if (jrec) {
            const uint32_t sn = so + P4, si = (uint32_t)i * Pid;
            hxjm = bld(rF, tjm, so + C(HX));
            hzjm = bld(rF, tjm, so + C(HZ));
            ezjm = bld(rF, tjm, so + C(EZ));
            exjm = bld(rF, tjm, so + C(EX));
            eyjm = bld(rF, tjm, so + C(EY));
            eyjmk = bld(rF, tjm + 4u, so + C(EY));
            eyjm1 = bld(rF, tjm, sn + C(EY));
            mj3 = bld_id<IDT>(rI, tidjm, si + 3 * ci);
            mj5 = bld_id<IDT>(rI, tidjm, si + 5 * ci);
        }
        float hx0k = 0.f, hy0k = 0.f, ezk = 0.f, ezjk = 0.f, eyk = 0.f, exk = 0.f, ez1k = 0.f;
        uint32_t mk3 = 0, mk4 = 0;
        if (krec) {
            const uint32_t si = (uint32_t)i * Pid;
            hx0k = bld(rF, tkm, so + C(HX));
            hy0k = bld(rF, tkm, so + C(HY));
            ezk = bld(rF, tkm, so + C(EZ));
            ezjk = bld(rF, tkm + 4u * NZ1, so + C(EZ));
            eyk = bld(rF, tkm, so + C(EY));
            exk = bld(rF, tkm, so + C(EX));
            ez1k = bld(rF, tkm, so + P4 + C(EZ));
            mk3 = bld_id<IDT>(rI, tidkm, si + 3 * ci);
            mk4 = bld_id<IDT>(rI, tidkm, si + 4 * ci);
        }

        // new H of the cell
        float hx, hy, hz;
        {
            const float4 qx = cH[cur.m3], qy = cH[cur.m4], qz = cH[cur.m5];
            hx = qx.x * cur.hx - qx.z * (cur.ezj - ez) + qx.w * (cur.eyk - ey);
            hy = qy.x * cur.hy - qy.w * (cur.exk - ex) + qy.y * (cur.ez1 - ez);
            hz = qz.x * cur.hz - qz.y * (cur.ey1 - ey) + qz.z * (cur.exj - ex);
        }
        if (pf & PF_SRCH) {
            hx = int_msrc(a, sH, i, j, k, 0, cur.m3, hx, it);
            hy = int_msrc(a, sH, i, j, k, 1, cur.m4, hy, it);
            hz = int_msrc(a, sH, i, j, k, 2, cur.m5, hz, it);
        }
        const int pb = (i - xa) & 1;
        lh[pb][0][tl] = hx;
        lh[pb][1][tl] = hy;
        lh[pb][2][tl] = hz;
        __syncthreads();
        // new Hx and Hz at j-1: from the lane Kin back, or recomputed
        float hxj, hzj;
        if (!jrec) {
            hxj = lh[pb][0][tl - Kin];
            hzj = lh[pb][2][tl - Kin];
        } else {
            const float4 qx = cH[mj3], qz = cH[mj5];
            hxj = qx.x * hxjm - qx.z * (ez - ezjm) + qx.w * (eyjmk - eyjm);
            hzj = qz.x * hzjm - qz.y * (eyjm1 - eyjm) + qz.z * (ex - exjm);
            if (pf & PF_SRCH) {
                hxj = int_msrc(a, sH, i, j - 1, k, 0, mj3, hxj, it);
                hzj = int_msrc(a, sH, i, j - 1, k, 2, mj5, hzj, it);
            }
        }
        // new Hx and Hy at k-1: from the previous lane, or recomputed
        float hxk, hyk;
        if (!krec) {
            hxk = lh[pb][0][tl - 1];
            hyk = lh[pb][1][tl - 1];
        } else {
            const float4 qx = cH[mk3], qy = cH[mk4];
            hxk = qx.x * hx0k - qx.z * (ezjk - ezk) + qx.w * (ey - eyk);
            hyk = qy.x * hy0k - qy.w * (ex - exk) + qy.y * (ez1k - ezk);
            if (pf & PF_SRCH) {
                hxk = int_msrc(a, sH, i, j, k - 1, 0, mk3, hxk, it);
                hyk = int_msrc(a, sH, i, j, k - 1, 1, mk4, hyk, it);
            }
        }

        if (live && (pf & PF_RX)) {  // the state at the start of the iteration
            for (int r = 0; r < a.nrx; ++r) {
                if (a.rxc[3 * r] != i || a.rxc[3 * r + 1] != j || a.rxc[3 * r + 2] != k) continue;
                float *o = a.rxbuf + (size_t)it * a.nrx + r;
                const size_t rs = (size_t)a.iterations * a.nrx;
                o[0] = ex; o[rs] = ey; o[2 * rs] = ez;
                o[3 * rs] = cur.hx; o[4 * rs] = cur.hy; o[5 * rs] = cur.hz;
            }
        }

        // new E of the cell
        float nex, ney, nez;
        {
            const float4 qx = cE[cur.m0], qy = cE[cur.m1], qz = cE[cur.m2];
            nex = qx.x * ex + qx.z * (hz - hzj) - qx.w * (hy - hyk);
            ney = qy.x * ey + qy.w * (hx - hxk) - qy.y * (hz - hzm);
            nez = qz.x * ez + qz.y * (hy - hym) - qz.z * (hx - hxj);
        }
        if (pf & PF_SRCE) {
            // Voltage sources, then Hertzian dipoles (gprMax/model_build_run.py:458-461).
            for (int s = 0; s < a.srcV.n; ++s) {  // gprMax/sources.py:81-120
                const int *in = a.srcV.info + 6 * s;
                if (in[0] != i || in[1] != j || in[2] != k || it < in[4] || it > in[5]) continue;
                const int pol = in[3] & 3;
                const bool hard = in[3] >= 4;
                const uint32_t m = pol == 0 ? cur.m0 : (pol == 1 ? cur.m1 : cur.m2);
                const float w = a.srcV.waves[(size_t)s * a.iterations + it];
                float *F = pol == 0 ? &nex : (pol == 1 ? &ney : &nez);
                if (hard) *F = -w / a.srcV.fac[2 * s];
                else *F = *F - sE[m] * w * a.srcV.fac[2 * s];
            }
            for (int s = 0; s < a.srcD.n; ++s) {  // gprMax/sources.py:163-193
                const int *in = a.srcD.info + 6 * s;
                if (in[0] != i || in[1] != j || in[2] != k || it < in[4] || it > in[5]) continue;
                const int pol = in[3];
                const uint32_t m = pol == 0 ? cur.m0 : (pol == 1 ? cur.m1 : cur.m2);
                const float w = a.srcD.waves[(size_t)s * a.iterations + it];
                const float v = sE[m] * w * a.srcD.fac[2 * s] * a.srcD.fac[2 * s + 1];
                if (pol == 0) nex = nex - v;
                else if (pol == 1) ney = ney - v;
                else nez = nez - v;
            }
        }
        if (live) {
            bst(wF, to, so + C(HX), hx);
            bst(wF, to, so + C(HY), hy);
            bst(wF, to, so + C(HZ), hz);
            bst(wF, to, so + C(EX), nex);
            bst(wF, to, so + C(EY), ney);
            bst(wF, to, so + C(EZ), nez);
        }
        hym = hy;
        hzm = hz;
        ex = cur.ex1;
        ey = cur.ey1;
        ez = cur.ez1;
    }
}

template <class IDT>
__global__ void __launch_bounds__(INT_BLOCK) k_interior(IntArgs a)
{
    const int item = logical_block(a.per_xcd);
    if (item >= a.nitems) return;
    interior_run<IDT>(a, item);
}

}  // namespace gpx
