/*
 * CPU oracle for the FDTD solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, loop for loop, the reference's Cython/OpenMP kernels of
 * the time-stepping hot path.  Only tests, the smoke check and the CPU-baseline
 * leg of the benchmark may load it, and only as the checker or as the timed
 * CPU baseline; the HIP product path never calls into it.
 *
 * It is compiled twice, with REAL=float (prefix orc_f32_) and REAL=double
 * (prefix orc_f64_), with -ffp-contract=off so that every expression rounds
 * exactly as written: (a*b)+c is never fused.  Expression trees keep the
 * reference's left-to-right association.
 *
 * Parallelism follows the reference: OpenMP static schedule over the outer
 * loop (the `prange(..., schedule='static')` of every Cython kernel).
 *
 * Array conventions (gprMax/grid.py:157-190):
 *   fields   [nx+1][ny+1][nz+1], z fastest
 *   ID       uint32 [6][nx+1][ny+1][nz+1]
 *   coeffs   [nmat][5]
 *   disp     [nmat][3*P][2]  (complex, re/im interleaved)
 *   T        [P][nx+1][ny+1][nz+1][2]
 *   PML Phi  [order][A][B][C] with the reference's padded shapes
 *            (gprMax/pml.py:202-219); R arrays [order][thickness]
 */
#include <stddef.h>
#include <stdint.h>

#ifndef REAL
#define REAL float
#endif
#ifndef ORC_PREFIX
#define ORC_PREFIX orc_f32
#endif
#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(ORC_PREFIX, name)

typedef REAL real;

/* Field index with nx+1, ny+1, nz+1 points per axis. */
#define F3(i, j, k) (((size_t)(i) * (size_t)(NY1) + (size_t)(j)) * (size_t)(NZ1) + (size_t)(k))
#define IDX(c, i, j, k) ((size_t)(c) * NTOT + F3(i, j, k))
#define CO(m, c) coeffs[(size_t)(m) * 5 + (c)]

/* ------------------------------------------------------------------------
 * update_electric (gprMax/fields_updates_ext.pyx:30-107).  A grid one cell
 * thick along an axis is a 2-D model (TMx, TMy, TMz) and updates only the
 * E component along that axis, over the ranges of lines 55-77; otherwise the
 * 3-D branch of lines 79-107.
 * ---------------------------------------------------------------------- */
void FN(update_electric)(int nx, int ny, int nz, int nthreads, const real *coeffs,
                         const uint32_t *ID, real *Ex, real *Ey, real *Ez,
                         const real *Hx, const real *Hy, const real *Hz)
{
    const long NY1 = ny + 1, NZ1 = nz + 1;
    const size_t NTOT = (size_t)(nx + 1) * NY1 * NZ1;
    long i, j, k;

    if (nx == 1) {          /* TMx: Ex, i in [0, nx), j in [1, ny), k in [1, nz) */
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
        for (i = 0; i < nx; i++)
            for (j = 1; j < ny; j++)
                for (k = 1; k < nz; k++) {
                    uint32_t m = ID[IDX(0, i, j, k)];
                    Ex[F3(i, j, k)] = CO(m, 0) * Ex[F3(i, j, k)] + CO(m, 2) * (Hz[F3(i, j, k)] - Hz[F3(i, j - 1, k)])
                                    - CO(m, 3) * (Hy[F3(i, j, k)] - Hy[F3(i, j, k - 1)]);
                }
        return;
    }
    if (ny == 1) {          /* TMy: Ey, i in [1, nx), j in [0, ny), k in [1, nz) */
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
        for (i = 1; i < nx; i++)
            for (j = 0; j < ny; j++)
                for (k = 1; k < nz; k++) {
                    uint32_t m = ID[IDX(1, i, j, k)];
                    Ey[F3(i, j, k)] = CO(m, 0) * Ey[F3(i, j, k)] + CO(m, 3) * (Hx[F3(i, j, k)] - Hx[F3(i, j, k - 1)])
                                    - CO(m, 1) * (Hz[F3(i, j, k)] - Hz[F3(i - 1, j, k)]);
                }
        return;
    }
    if (nz == 1) {          /* TMz: Ez, i in [1, nx), j in [1, ny), k in [0, nz) */
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
        for (i = 1; i < nx; i++)
            for (j = 1; j < ny; j++)
                for (k = 0; k < nz; k++) {
                    uint32_t m = ID[IDX(2, i, j, k)];
                    Ez[F3(i, j, k)] = CO(m, 0) * Ez[F3(i, j, k)] + CO(m, 1) * (Hy[F3(i, j, k)] - Hy[F3(i - 1, j, k)])
                                    - CO(m, 2) * (Hx[F3(i, j, k)] - Hx[F3(i, j - 1, k)]);
                }
        return;
    }

#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 1; i < nx; i++)
        for (j = 1; j < ny; j++)
            for (k = 1; k < nz; k++) {
                uint32_t mx = ID[IDX(0, i, j, k)], my = ID[IDX(1, i, j, k)], mz = ID[IDX(2, i, j, k)];
                Ex[F3(i, j, k)] = CO(mx, 0) * Ex[F3(i, j, k)] + CO(mx, 2) * (Hz[F3(i, j, k)] - Hz[F3(i, j - 1, k)])
                                - CO(mx, 3) * (Hy[F3(i, j, k)] - Hy[F3(i, j, k - 1)]);
                Ey[F3(i, j, k)] = CO(my, 0) * Ey[F3(i, j, k)] + CO(my, 3) * (Hx[F3(i, j, k)] - Hx[F3(i, j, k - 1)])
                                - CO(my, 1) * (Hz[F3(i, j, k)] - Hz[F3(i - 1, j, k)]);
                Ez[F3(i, j, k)] = CO(mz, 0) * Ez[F3(i, j, k)] + CO(mz, 1) * (Hy[F3(i, j, k)] - Hy[F3(i - 1, j, k)])
                                - CO(mz, 2) * (Hx[F3(i, j, k)] - Hx[F3(i, j - 1, k)]);
            }

    /* Ex at i = 0 */
#pragma omp parallel for schedule(static) num_threads(nthreads) private(k)
    for (j = 1; j < ny; j++)
        for (k = 1; k < nz; k++) {
            uint32_t m = ID[IDX(0, 0, j, k)];
            Ex[F3(0, j, k)] = CO(m, 0) * Ex[F3(0, j, k)] + CO(m, 2) * (Hz[F3(0, j, k)] - Hz[F3(0, j - 1, k)])
                            - CO(m, 3) * (Hy[F3(0, j, k)] - Hy[F3(0, j, k - 1)]);
        }
    /* Ey at j = 0 */
#pragma omp parallel for schedule(static) num_threads(nthreads) private(k)
    for (i = 1; i < nx; i++)
        for (k = 1; k < nz; k++) {
            uint32_t m = ID[IDX(1, i, 0, k)];
            Ey[F3(i, 0, k)] = CO(m, 0) * Ey[F3(i, 0, k)] + CO(m, 3) * (Hx[F3(i, 0, k)] - Hx[F3(i, 0, k - 1)])
                            - CO(m, 1) * (Hz[F3(i, 0, k)] - Hz[F3(i - 1, 0, k)]);
        }
    /* Ez at k = 0 */
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j)
    for (i = 1; i < nx; i++)
        for (j = 1; j < ny; j++) {
            uint32_t m = ID[IDX(2, i, j, 0)];
            Ez[F3(i, j, 0)] = CO(m, 0) * Ez[F3(i, j, 0)] + CO(m, 1) * (Hy[F3(i, j, 0)] - Hy[F3(i - 1, j, 0)])
                            - CO(m, 2) * (Hx[F3(i, j, 0)] - Hx[F3(i, j - 1, 0)]);
        }
}

/* ------------------------------------------------------------------------
 * update_magnetic (gprMax/fields_updates_ext.pyx:352-412).  A 2-D model
 * updates Hx when ny or nz is 1, Hy when nx or nz is 1 and Hz when nx or ny
 * is 1, each with the coefficient and field at the same (i, j, k) and the
 * ranges of lines 376-400; otherwise the 3-D branch (lines 402-412).
 * ---------------------------------------------------------------------- */
void FN(update_magnetic)(int nx, int ny, int nz, int nthreads, const real *coeffs,
                         const uint32_t *ID, const real *Ex, const real *Ey, const real *Ez,
                         real *Hx, real *Hy, real *Hz)
{
    const long NY1 = ny + 1, NZ1 = nz + 1;
    const size_t NTOT = (size_t)(nx + 1) * NY1 * NZ1;
    long i, j, k;

    if (nx == 1 || ny == 1 || nz == 1) {
        if (ny == 1 || nz == 1) {
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
            for (i = 1; i < nx; i++)
                for (j = 0; j < ny; j++)
                    for (k = 0; k < nz; k++) {
                        uint32_t m = ID[IDX(3, i, j, k)];
                        Hx[F3(i, j, k)] = CO(m, 0) * Hx[F3(i, j, k)] - CO(m, 2) * (Ez[F3(i, j + 1, k)] - Ez[F3(i, j, k)])
                                        + CO(m, 3) * (Ey[F3(i, j, k + 1)] - Ey[F3(i, j, k)]);
                    }
        }
        if (nx == 1 || nz == 1) {
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
            for (i = 0; i < nx; i++)
                for (j = 1; j < ny; j++)
                    for (k = 0; k < nz; k++) {
                        uint32_t m = ID[IDX(4, i, j, k)];
                        Hy[F3(i, j, k)] = CO(m, 0) * Hy[F3(i, j, k)] - CO(m, 3) * (Ex[F3(i, j, k + 1)] - Ex[F3(i, j, k)])
                                        + CO(m, 1) * (Ez[F3(i + 1, j, k)] - Ez[F3(i, j, k)]);
                    }
        }
        if (nx == 1 || ny == 1) {
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
            for (i = 0; i < nx; i++)
                for (j = 0; j < ny; j++)
                    for (k = 1; k < nz; k++) {
                        uint32_t m = ID[IDX(5, i, j, k)];
                        Hz[F3(i, j, k)] = CO(m, 0) * Hz[F3(i, j, k)] - CO(m, 1) * (Ey[F3(i + 1, j, k)] - Ey[F3(i, j, k)])
                                        + CO(m, 2) * (Ex[F3(i, j + 1, k)] - Ex[F3(i, j, k)]);
                    }
        }
        return;
    }

#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 0; i < nx; i++)
        for (j = 0; j < ny; j++)
            for (k = 0; k < nz; k++) {
                uint32_t mx = ID[IDX(3, i + 1, j, k)], my = ID[IDX(4, i, j + 1, k)], mz = ID[IDX(5, i, j, k + 1)];
                Hx[F3(i + 1, j, k)] = CO(mx, 0) * Hx[F3(i + 1, j, k)]
                                    - CO(mx, 2) * (Ez[F3(i + 1, j + 1, k)] - Ez[F3(i + 1, j, k)])
                                    + CO(mx, 3) * (Ey[F3(i + 1, j, k + 1)] - Ey[F3(i + 1, j, k)]);
                Hy[F3(i, j + 1, k)] = CO(my, 0) * Hy[F3(i, j + 1, k)]
                                    - CO(my, 3) * (Ex[F3(i, j + 1, k + 1)] - Ex[F3(i, j + 1, k)])
                                    + CO(my, 1) * (Ez[F3(i + 1, j + 1, k)] - Ez[F3(i, j + 1, k)]);
                Hz[F3(i, j, k + 1)] = CO(mz, 0) * Hz[F3(i, j, k + 1)]
                                    - CO(mz, 1) * (Ey[F3(i + 1, j, k + 1)] - Ey[F3(i, j, k + 1)])
                                    + CO(mz, 2) * (Ex[F3(i, j + 1, k + 1)] - Ex[F3(i, j, k + 1)]);
            }
}

/* ------------------------------------------------------------------------
 * Dispersive updates, 3-D (gprMax/fields_updates_ext.pyx:113-346).
 * T and the coefficients are complex; the products follow complex
 * arithmetic term by term: (a+ib)(c+id) = (ac-bd) + i(ad+bc), and a
 * complex times a real field value scales both parts.
 * One-pole form (maxpoles == 1) starts phi from the first product, the
 * multipole form from zero, as the two reference functions do.
 * ---------------------------------------------------------------------- */
#define DRE(m, n) disp[((size_t)(m) * 3 * P + (n)) * 2]
#define DIM(m, n) disp[((size_t)(m) * 3 * P + (n)) * 2 + 1]
#define TRE(T, p, f) T[((size_t)(p) * NTOT + (f)) * 2]
#define TIM(T, p, f) T[((size_t)(p) * NTOT + (f)) * 2 + 1]

static inline real disp_A_cell(int P, const real *disp, uint32_t m, real *T, size_t NTOT, size_t f, real E)
{
    real phi = 0;
    for (int p = 0; p < P; p++) {
        real tr = TRE(T, p, f), ti = TIM(T, p, f);
        if (P == 1)
            phi = DRE(m, 0) * tr;
        else
            phi = phi + DRE(m, 3 * p) * tr;
        real er = DRE(m, 3 * p + 1), ei = DIM(m, 3 * p + 1);
        real zr = DRE(m, 3 * p + 2), zi = DIM(m, 3 * p + 2);
        real nr = (er * tr - ei * ti) + zr * E;
        real ni = (er * ti + ei * tr) + zi * E;
        TRE(T, p, f) = nr;
        TIM(T, p, f) = ni;
    }
    return phi;
}

void FN(update_electric_dispersive_A)(int nx, int ny, int nz, int nthreads, int P,
                                      const real *coeffs, const real *disp, const uint32_t *ID,
                                      real *Tx, real *Ty, real *Tz,
                                      real *Ex, real *Ey, real *Ez,
                                      const real *Hx, const real *Hy, const real *Hz)
{
    const long NY1 = ny + 1, NZ1 = nz + 1;
    const size_t NTOT = (size_t)(nx + 1) * NY1 * NZ1;
    long i, j, k;
    /* Each component is skipped when its two transverse axes are both one
     * cell (fields_updates_ext.pyx:146, 158, 170). */
    /* Ex: i in [0, nx), j, k in [1, n) */
    if (ny != 1 || nz != 1) {
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 0; i < nx; i++)
        for (j = 1; j < ny; j++)
            for (k = 1; k < nz; k++) {
                size_t f = F3(i, j, k);
                uint32_t m = ID[IDX(0, i, j, k)];
                real phi = disp_A_cell(P, disp, m, Tx, NTOT, f, Ex[f]);
                Ex[f] = CO(m, 0) * Ex[f] + CO(m, 2) * (Hz[f] - Hz[F3(i, j - 1, k)])
                      - CO(m, 3) * (Hy[f] - Hy[F3(i, j, k - 1)]) - CO(m, 4) * phi;
            }
    }
    /* Ey: i in [1, nx), j in [0, ny), k in [1, nz) */
    if (nx != 1 || nz != 1) {
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 1; i < nx; i++)
        for (j = 0; j < ny; j++)
            for (k = 1; k < nz; k++) {
                size_t f = F3(i, j, k);
                uint32_t m = ID[IDX(1, i, j, k)];
                real phi = disp_A_cell(P, disp, m, Ty, NTOT, f, Ey[f]);
                Ey[f] = CO(m, 0) * Ey[f] + CO(m, 3) * (Hx[f] - Hx[F3(i, j, k - 1)])
                      - CO(m, 1) * (Hz[f] - Hz[F3(i - 1, j, k)]) - CO(m, 4) * phi;
            }
    }
    /* Ez: i, j in [1, n), k in [0, nz) */
    if (nx != 1 || ny != 1) {
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 1; i < nx; i++)
        for (j = 1; j < ny; j++)
            for (k = 0; k < nz; k++) {
                size_t f = F3(i, j, k);
                uint32_t m = ID[IDX(2, i, j, k)];
                real phi = disp_A_cell(P, disp, m, Tz, NTOT, f, Ez[f]);
                Ez[f] = CO(m, 0) * Ez[f] + CO(m, 1) * (Hy[f] - Hy[F3(i - 1, j, k)])
                      - CO(m, 2) * (Hx[f] - Hx[F3(i, j - 1, k)]) - CO(m, 4) * phi;
            }
    }
}

static inline void disp_B_cell(int P, const real *disp, uint32_t m, real *T, size_t NTOT, size_t f, real E)
{
    for (int p = 0; p < P; p++) {
        TRE(T, p, f) = TRE(T, p, f) - DRE(m, 3 * p + 2) * E;
        TIM(T, p, f) = TIM(T, p, f) - DIM(m, 3 * p + 2) * E;
    }
}

void FN(update_electric_dispersive_B)(int nx, int ny, int nz, int nthreads, int P,
                                      const real *disp, const uint32_t *ID,
                                      real *Tx, real *Ty, real *Tz,
                                      const real *Ex, const real *Ey, const real *Ez)
{
    const long NY1 = ny + 1, NZ1 = nz + 1;
    const size_t NTOT = (size_t)(nx + 1) * NY1 * NZ1;
    long i, j, k;
    /* Guards of fields_updates_ext.pyx:211, 220, 229 */
    if (ny != 1 || nz != 1)
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 0; i < nx; i++)
        for (j = 1; j < ny; j++)
            for (k = 1; k < nz; k++)
                disp_B_cell(P, disp, ID[IDX(0, i, j, k)], Tx, NTOT, F3(i, j, k), Ex[F3(i, j, k)]);
    if (nx != 1 || nz != 1)
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 1; i < nx; i++)
        for (j = 0; j < ny; j++)
            for (k = 1; k < nz; k++)
                disp_B_cell(P, disp, ID[IDX(1, i, j, k)], Ty, NTOT, F3(i, j, k), Ey[F3(i, j, k)]);
    if (nx != 1 || ny != 1)
#pragma omp parallel for schedule(static) num_threads(nthreads) private(j, k)
    for (i = 1; i < nx; i++)
        for (j = 1; j < ny; j++)
            for (k = 0; k < nz; k++)
                disp_B_cell(P, disp, ID[IDX(2, i, j, k)], Tz, NTOT, F3(i, j, k), Ez[F3(i, j, k)]);
}

/* ------------------------------------------------------------------------
 * PML slab corrections, HORIPML and MRIPML, orders 1 and 2, six directions
 * (gprMax/pml_updates/pml_updates_{electric,magnetic}_{HORIPML,MRIPML}_ext.pyx).
 *
 * direction: 0 xminus, 1 yminus, 2 zminus, 3 xplus, 4 yplus, 5 zplus.
 * Loop indices (a, b, c) run over the slab box; the index along the slab
 * normal is the profile index p.  Along the normal the cell is
 *   E, minus slab:  f - p          E, plus slab: s + p
 *   H, minus slab:  f - (p + 1)    H, plus slab: s + p
 * Corrected components, sign and source difference per axis:
 *   E x: Ey -= (Hz, Phi1), Ez += (Hy, Phi2)   backward difference / d
 *   E y: Ex += (Hz, Phi1), Ez -= (Hx, Phi2)
 *   E z: Ex -= (Hy, Phi1), Ey += (Hx, Phi2)
 *   H x: Hy += (Ez, Phi1), Hz -= (Ey, Phi2)   forward difference / d
 *   H y: Hx -= (Ez, Phi1), Hz += (Ex, Phi2)
 *   H z: Hx += (Ey, Phi1), Hy -= (Ex, Phi2)
 * ---------------------------------------------------------------------- */
struct pml_coef {
    real RA0, RA1, RB0, RB1, RE0, RE1, RF0, RF1;
    real RA01;            /* HORIPML */
    real IRA, IRA1, RC0, RC1; /* MRIPML */
};

static inline struct pml_coef pml_coef_at(int formulation, int order, int t, int p,
                                          const real *RA, const real *RB, const real *RE, const real *RF)
{
    struct pml_coef q = {0};
    q.RA0 = RA[p];
    q.RB0 = RB[p];
    q.RE0 = RE[p];
    q.RF0 = RF[p];
    if (order == 2) {
        q.RA1 = RA[t + p];
        q.RB1 = RB[t + p];
        q.RE1 = RE[t + p];
        q.RF1 = RF[t + p];
    }
    if (formulation == 0) {
        if (order == 1)
            q.RA01 = RA[p] - 1;
        else
            q.RA01 = RA[p] * RA[t + p] - 1;
    } else {
        if (order == 1) {
            q.IRA = 1 / RA[p];
            q.IRA1 = q.IRA - 1;
            q.RC0 = q.IRA * q.RB0 * q.RF0;
        } else {
            q.IRA = 1 / (RA[p] + RA[t + p]);
            q.IRA1 = q.IRA - 1;
            q.RC0 = q.IRA * q.RF0;
            q.RC1 = q.IRA * q.RF1;
        }
    }
    return q;
}

/* One corrected component value; phi points at order-0 Phi, phi_o at order 1. */
static inline real pml_apply(int formulation, int order, const struct pml_coef *q, real F, int sign,
                             real c, real d, real *phi, real *phi_o)
{
    real corr;
    if (formulation == 0) {
        if (order == 1) {
            corr = c * (q->RA01 * d + q->RB0 * *phi);
            *phi = q->RE0 * *phi - q->RF0 * d;
        } else {
            corr = c * (q->RA01 * d + q->RA1 * q->RB0 * *phi + q->RB1 * *phi_o);
            *phi_o = q->RE1 * *phi_o - q->RF1 * (q->RA0 * d + q->RB0 * *phi);
            *phi = q->RE0 * *phi - q->RF0 * d;
        }
    } else {
        if (order == 1) {
            corr = c * (q->IRA1 * d - q->IRA * *phi);
            *phi = q->RE0 * *phi + q->RC0 * d - q->RC0 * *phi;
        } else {
            real psi = q->RB0 * *phi + q->RB1 * *phi_o;
            corr = c * (q->IRA1 * d - q->IRA * psi);
            *phi_o = q->RE1 * *phi_o + q->RC1 * (d - psi);
            *phi = q->RE0 * *phi + q->RC0 * (d - psi);
        }
    }
    return sign > 0 ? F + corr : F - corr;
}

static void phi_dims(int field, int axis, int snx, int sny, int snz, long dims1[3], long dims2[3])
{
    /* gprMax/pml.py:202-219 */
    if (field == 0) {
        if (axis == 0) { long a[3] = {snx + 1, sny, snz + 1}, b[3] = {snx + 1, sny + 1, snz}; for (int n = 0; n < 3; n++) { dims1[n] = a[n]; dims2[n] = b[n]; } }
        else if (axis == 1) { long a[3] = {snx, sny + 1, snz + 1}, b[3] = {snx + 1, sny + 1, snz}; for (int n = 0; n < 3; n++) { dims1[n] = a[n]; dims2[n] = b[n]; } }
        else { long a[3] = {snx, sny + 1, snz + 1}, b[3] = {snx + 1, sny, snz + 1}; for (int n = 0; n < 3; n++) { dims1[n] = a[n]; dims2[n] = b[n]; } }
    } else {
        if (axis == 0) { long a[3] = {snx, sny + 1, snz}, b[3] = {snx, sny, snz + 1}; for (int n = 0; n < 3; n++) { dims1[n] = a[n]; dims2[n] = b[n]; } }
        else if (axis == 1) { long a[3] = {snx + 1, sny, snz}, b[3] = {snx, sny, snz + 1}; for (int n = 0; n < 3; n++) { dims1[n] = a[n]; dims2[n] = b[n]; } }
        else { long a[3] = {snx + 1, sny, snz}, b[3] = {snx, sny + 1, snz}; for (int n = 0; n < 3; n++) { dims1[n] = a[n]; dims2[n] = b[n]; } }
    }
}

void FN(pml_update)(int field, int formulation, int order, int direction,
                    int xs, int xf, int ys, int yf, int zs, int zf, int nthreads,
                    int nx, int ny, int nz, const real *coeffs, const uint32_t *ID,
                    real *Ex, real *Ey, real *Ez, real *Hx, real *Hy, real *Hz,
                    real *Phi1, real *Phi2, const real *RA, const real *RB,
                    const real *RE, const real *RF, real d)
{
    const long NY1 = ny + 1, NZ1 = nz + 1;
    const size_t NTOT = (size_t)(nx + 1) * NY1 * NZ1;
    const int axis = direction % 3, minus = direction < 3;
    const int snx = xf - xs, sny = yf - ys, snz = zf - zs;
    const int t = axis == 0 ? snx : (axis == 1 ? sny : snz);
    long dims1[3], dims2[3];
    phi_dims(field, axis, snx, sny, snz, dims1, dims2);
    const size_t o1 = (size_t)dims1[0] * dims1[1] * dims1[2], o2 = (size_t)dims2[0] * dims2[1] * dims2[2];

    /* component tables: corrected array, its ID row, source array, sign */
    real *F1, *F2;
    const real *S1, *S2;
    int c1, c2, s1, s2;
    if (field == 0) {
        if (axis == 0) { F1 = Ey; c1 = 1; S1 = Hz; s1 = -1; F2 = Ez; c2 = 2; S2 = Hy; s2 = +1; }
        else if (axis == 1) { F1 = Ex; c1 = 0; S1 = Hz; s1 = +1; F2 = Ez; c2 = 2; S2 = Hx; s2 = -1; }
        else { F1 = Ex; c1 = 0; S1 = Hy; s1 = -1; F2 = Ey; c2 = 1; S2 = Hx; s2 = +1; }
    } else {
        if (axis == 0) { F1 = Hy; c1 = 4; S1 = Ez; s1 = +1; F2 = Hz; c2 = 5; S2 = Ey; s2 = -1; }
        else if (axis == 1) { F1 = Hx; c1 = 3; S1 = Ez; s1 = -1; F2 = Hz; c2 = 5; S2 = Ex; s2 = +1; }
        else { F1 = Hx; c1 = 3; S1 = Ey; s1 = +1; F2 = Hy; c2 = 4; S2 = Ex; s2 = -1; }
    }
    const size_t step = axis == 0 ? (size_t)NY1 * NZ1 : (axis == 1 ? (size_t)NZ1 : 1);

    long a;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (a = 0; a < snx; a++)
        for (long b = 0; b < sny; b++)
            for (long c = 0; c < snz; c++) {
                long idx[3] = {a, b, c};
                long p = idx[axis];
                long cell[3] = {xs + a, ys + b, zs + c};
                long fend = axis == 0 ? xf : (axis == 1 ? yf : zf);
                if (minus)
                    cell[axis] = field == 0 ? fend - p : fend - (p + 1);
                struct pml_coef q = pml_coef_at(formulation, order, t, (int)p, RA, RB, RE, RF);
                size_t f = F3(cell[0], cell[1], cell[2]);
                size_t ph1 = ((size_t)a * dims1[1] + b) * dims1[2] + c;
                size_t ph2 = ((size_t)a * dims2[1] + b) * dims2[2] + c;
                real dd1, dd2;
                if (field == 0) {
                    dd1 = (S1[f] - S1[f - step]) / d;
                } else {
                    dd1 = (S1[f + step] - S1[f]) / d;
                }
                F1[f] = pml_apply(formulation, order, &q, F1[f], s1, CO(ID[(size_t)c1 * NTOT + f], 4), dd1,
                                  &Phi1[ph1], order == 2 ? &Phi1[o1 + ph1] : 0);
                if (field == 0) {
                    dd2 = (S2[f] - S2[f - step]) / d;
                } else {
                    dd2 = (S2[f + step] - S2[f]) / d;
                }
                F2[f] = pml_apply(formulation, order, &q, F2[f], s2, CO(ID[(size_t)c2 * NTOT + f], 4), dd2,
                                  &Phi2[ph2], order == 2 ? &Phi2[o2 + ph2] : 0);
            }
}

/* ------------------------------------------------------------------------
 * Snapshot cell averages (gprMax/snapshots_ext.pyx:24-71 on the slices of
 * gprMax/snapshots.py:78-100).  Output arrays are [snx][sny][snz].
 * ---------------------------------------------------------------------- */
void FN(snapshot_fields)(int nx, int ny, int nz, int xs, int ys, int zs, int sdx, int sdy, int sdz,
                         int snx, int sny, int snz,
                         const real *Ex, const real *Ey, const real *Ez,
                         const real *Hx, const real *Hy, const real *Hz,
                         real *oEx, real *oEy, real *oEz, real *oHx, real *oHy, real *oHz)
{
    const long NY1 = ny + 1, NZ1 = nz + 1;
    (void)nx;
#define SL(i, j, k) F3(xs + (long)(i) * sdx, ys + (long)(j) * sdy, zs + (long)(k) * sdz)
    for (long i = 0; i < snx; i++)
        for (long j = 0; j < sny; j++)
            for (long k = 0; k < snz; k++) {
                size_t o = ((size_t)i * sny + j) * snz + k;
                oEx[o] = (Ex[SL(i, j, k)] + Ex[SL(i, j + 1, k)] + Ex[SL(i, j, k + 1)] + Ex[SL(i, j + 1, k + 1)]) / 4;
                oEy[o] = (Ey[SL(i, j, k)] + Ey[SL(i + 1, j, k)] + Ey[SL(i, j, k + 1)] + Ey[SL(i + 1, j, k + 1)]) / 4;
                oEz[o] = (Ez[SL(i, j, k)] + Ez[SL(i + 1, j, k)] + Ez[SL(i, j + 1, k)] + Ez[SL(i + 1, j + 1, k)]) / 4;
                oHx[o] = (Hx[SL(i, j, k)] + Hx[SL(i + 1, j, k)]) / 2;
                oHy[o] = (Hy[SL(i, j, k)] + Hy[SL(i, j + 1, k)]) / 2;
                oHz[o] = (Hz[SL(i, j, k)] + Hz[SL(i, j, k + 1)]) / 2;
            }
#undef SL
}
