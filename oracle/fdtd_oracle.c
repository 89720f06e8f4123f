/*
 * fdtd_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A CPU, FP64, single-threaded restatement of the split-field FDTD time-stepper of
 * SMatrAnEM, FDTDcalc/FDTDsplitfieldMetal.c (cited below as F:<line>).  It exists to
 * check the HIP library; it is never linked into, loaded by or called from the
 * product.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off (oracle/Makefile).  With contraction off
 * and the same operation order as the reference, every FP64 result is bit-identical to
 * the reference compiled with gcc -O2 on x86-64 (no FMA without -march).
 *
 * Storage contract (F:151-152, F:188-233): fields H_t, E_t are double[3*(n+Lnotn+1)],
 * component-interleaved main slots X[c+3l] for l < n = L0*L1*L2, then the split/aux
 * slots X[3n + c + 3*addr[l]]; addr[l] == Lnotn names the never-written zero slot.
 * Complex arrays are interleaved (re, im) doubles, the memory layout of C99
 * double complex.
 *
 * Pinning of each function (the tests that check it):
 *   orc_main_calc        bitwise vs the compiled reference main_calc (F:242-413) on
 *                        random states with every cell class; brute-force Yee curl on
 *                        3x3x3; adjointness of the two curls; zero-source invariance.
 *   orc_set_env          bitwise vs reference set_env (F:452-475); sum over a run == 1.
 *   orc_set_tau          bitwise vs reference set_tau (F:515-524); exactness at omega.
 *   orc_expn_table       closed form cexp(i*phi) at the reference phases (F:93-99).
 *   orc_set_sigma        bitwise vs reference set_sigma (F:858-926); round trip of the
 *   orc_remove_sigma     PML stretch (PMLset/PMLcalc.c:230) through both functions.
 *   orc_set_address      bitwise vs reference set_address (F:188-233); brute-force counts.
 *   orc_build_coeffs     not run against the reference (the loop is inline in main,
 *                        F:106-143, which needs BLAS to run); pinned by closed forms
 *                        of every branch: one cell of each class, driven at omega through
 *                        orc_main_calc, settles to the frequency-domain response the
 *                        scheme is exact for -- lossy eps and lossy mu ('n'), the main
 *                        and split slots of both families in PML with distinct sigma
 *                        samples per axis ('a'), and the eps- and mu-metal stable forms
 *                        of check_files/StableForm.pdf p.3 ('e', 'h', 'b') -- and
 *                        A = 1, B = 1/(iR*eps) in lossless cells.
 *   orc_run              composition of the above in the order of F:153-167.
 *   orc_point_add,       no reference counterpart (a point current source on one E
 *   orc_run_point        component, the reference's source being a sheet): pinned by the
 *                        frequency-domain solve of a CW point source in a PML box
 *                        (tests/test_oracle_freqdomain.py), which fixes the cell, the
 *                        component, the coefficient and the step timing of the term.
 *   orc_s_project        bitwise vs reference S_calc (F:419-448) reading the same files.
 *   orc_step_adjoint,    no reference counterpart (this build's adjoint): the transpose
 *   orc_adjoint_run      is pinned by the dot-product test <M x, y> = <x, M^T y> and the
 *                        gradient by central finite differences of orc_run.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EMIN (10. * DBL_EPSILON)                              /* constant.h:7 */
#define ORC_2PI (2.L * 3.1415926535897932384626433832795028841971L) /* constant.h:15 */

/* ---------------------------------------------------------------------------------- */
/* Neighbours and curls (F:330-413)                                                    */
/* ---------------------------------------------------------------------------------- */

typedef struct {
    int64_t off[3]; /* signed cell offset of the neighbour along each axis */
    double w[3];    /* Bloch/PEC weight: iBc[d] across the wrap, else 1 */
} orc_nb;

/* Backward neighbours: j-1 with wrap, weight iBc at j == 0 (F:396-402). */
static orc_nb nb_backward(const int L[3], const int iBc[3], const int j[3])
{
    orc_nb nb;
    int64_t stride = 1;
    for (int d = 0; d < 3; d++) {
        int jm = (j[d] == 0) ? L[d] - 1 : j[d] - 1;
        nb.off[d] = (int64_t)(jm - j[d]) * stride;
        nb.w[d] = (j[d] == 0) ? (double)iBc[d] : 1.;
        stride *= L[d];
    }
    return nb;
}

/* Forward neighbours: j+1 with wrap, weight iBc at j == L-1 (F:375-381). */
static orc_nb nb_forward(const int L[3], const int iBc[3], const int j[3])
{
    orc_nb nb;
    int64_t stride = 1;
    for (int d = 0; d < 3; d++) {
        int jp = (j[d] + 1 == L[d]) ? 0 : j[d] + 1;
        nb.off[d] = (int64_t)(jp - j[d]) * stride;
        nb.w[d] = (j[d] + 1 == L[d]) ? (double)iBc[d] : 1.;
        stride *= L[d];
    }
    return nb;
}

/*
 * Slot lookup: with map == NULL the value of component c at cell q is P[c + 3q]
 * (main slots, ope_R/ope_Rd); otherwise P[c + 3*map[q]] (split slots through
 * address_notn, ope_Rmap/ope_Rdmap).
 */
static inline double slot(const double *P, const int *map, int c, int64_t q)
{
    return map ? P[c + 3 * (int64_t)map[q]] : P[c + 3 * q];
}

/*
 * Backward curl pieces (ope_Rd / ope_Rdmap, F:393-413, F:351-371).  half selects the
 * terms: 0 -> '+' only, 1 -> '-' only, 2 -> both, accumulated as (0 + T+) + T-.
 *   T+[c] =  P[c2](l) - w[c1]*P[c2](l - e_c1)
 *   T-[c] = -P[c1](l) + w[c2]*P[c1](l - e_c2)
 */
static void curl_backward(int half, const orc_nb *nb, const double *P, const int *map,
                          int64_t l, double W[3])
{
    for (int c = 0; c < 3; c++) {
        int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        double acc = 0.;
        if (half != 1) {
            double here = slot(P, map, c2, l);
            double there = slot(P, map, c2, l + nb->off[c1]);
            acc += here - nb->w[c1] * there;
        }
        if (half != 0) {
            double here = slot(P, map, c1, l);
            double there = slot(P, map, c1, l + nb->off[c2]);
            acc += -here + nb->w[c2] * there;
        }
        W[c] = acc;
    }
}

/*
 * Forward curl pieces (ope_R / ope_Rmap, F:372-392, F:330-350).
 *   T+[c] =  w[c1]*P[c2](l + e_c1) - P[c2](l)
 *   T-[c] = -w[c2]*P[c1](l + e_c2) + P[c1](l)
 */
static void curl_forward(int half, const orc_nb *nb, const double *P, const int *map,
                         int64_t l, double W[3])
{
    for (int c = 0; c < 3; c++) {
        int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        double acc = 0.;
        if (half != 1) {
            double there = slot(P, map, c2, l + nb->off[c1]);
            double here = slot(P, map, c2, l);
            acc += nb->w[c1] * there - here;
        }
        if (half != 0) {
            double there = slot(P, map, c1, l + nb->off[c2]);
            double here = slot(P, map, c1, l);
            acc += -nb->w[c2] * there + here;
        }
        W[c] = acc;
    }
}

/* ---------------------------------------------------------------------------------- */
/* One leapfrog step (main_calc, F:242-329)                                             */
/* ---------------------------------------------------------------------------------- */

/*
 * Same arguments and layouts as the reference main_calc (F:242-249); complex arrays
 * as interleaved doubles.  Returns 0, or 1 + m when monitor plane m holds a PML ('a')
 * cell (the reference prints "iBTIin[m] error!" and exits, F:320).
 */
int orc_main_calc(const int L[4], const double *Ah, const double *Bh, const double *Ae,
                  const double *Be, const int iBc[3], double *H_t, double *E_t,
                  const int iBTIin[4], const double *Hi, const double *Ei,
                  const double *ExpN2, double *HB, double *EB, double *HT, double *ET,
                  const char *Cont, const int *addr)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2];
    const int64_t n = L0 * L1 * L2;
    const int iI = iBTIin[2];
    const double sgn = (double)(iBTIin[3] / abs(iBTIin[3])); /* F:253 */
    double *Hs = H_t + 3 * n, *Es = E_t + 3 * n;          /* split region */

    /* H half step (F:255-286): reads E only, so the cell order is immaterial; the OpenMP
     * build of this file (liboracle_omp.so, a CPU baseline) runs the planes on several
     * threads with the same per-cell arithmetic. */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < (int)L2; z++) {
        double W[3];
        int j[3];
        j[2] = z;
        for (j[1] = 0; j[1] < L1; j[1]++)
            for (j[0] = 0; j[0] < L0; j[0]++) {
                const int64_t l = j[0] + L0 * j[1] + L0 * L1 * j[2];
                const char k = Cont[l];
                if (k == 'e') continue; /* Q3: H frozen in epsilon-metal cells */
                orc_nb nb = nb_backward(L, iBc, j);
                if (k == 'n') {
                    curl_backward(2, &nb, E_t, NULL, l, W);
                    if (j[2] == iI) { /* modal sheet source, 'n' cells only (F:261-264) */
                        const double *src = Ei + 2 * (j[0] + L0 * j[1]);
                        W[0] += sgn * src[1];
                        W[1] -= sgn * src[0];
                    }
                    for (int c = 0; c < 3; c++)
                        H_t[c + 3 * l] = Ah[c + 3 * l] * H_t[c + 3 * l] + Bh[c + 3 * l] * (-W[c]);
                    curl_backward(2, &nb, Es, addr, l, W);
                    for (int c = 0; c < 3; c++) H_t[c + 3 * l] += Bh[c + 3 * l] * (-W[c]);
                } else if (k == 'a') { /* Berenger split: '+' in main, '-' in split (F:268-276) */
                    for (int half = 0; half < 2; half++) {
                        const int64_t s = half == 0 ? 3 * l : 3 * n + 3 * (int64_t)addr[l];
                        curl_backward(half, &nb, E_t, NULL, l, W);
                        for (int c = 0; c < 3; c++)
                            H_t[s + c] = Ah[s + c] * H_t[s + c] + Bh[s + c] * (-W[c]);
                        curl_backward(half, &nb, Es, addr, l, W);
                        for (int c = 0; c < 3; c++) H_t[s + c] += Bh[s + c] * (-W[c]);
                    }
                } else { /* 'h' or 'b': stable form, aux in the split slot (F:277-283) */
                    const int64_t s = 3 * n + 3 * (int64_t)addr[l];
                    curl_backward(2, &nb, E_t, NULL, l, W);
                    for (int c = 0; c < 3; c++) H_t[c + 3 * l] += 0.5 * H_t[s + c];
                    for (int c = 0; c < 3; c++)
                        H_t[s + c] = Ah[s + c] * H_t[s + c] + Bh[s + c] * (-W[c]);
                    for (int c = 0; c < 3; c++) H_t[c + 3 * l] += 0.5 * H_t[s + c];
                }
            }
    }

    /* E half step (F:287-317): reads H only. */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < (int)L2; z++) {
        double W[3];
        int j[3];
        j[2] = z;
        for (j[1] = 0; j[1] < L1; j[1]++)
            for (j[0] = 0; j[0] < L0; j[0]++) {
                const int64_t l = j[0] + L0 * j[1] + L0 * L1 * j[2];
                const char k = Cont[l];
                if (k == 'h') continue; /* Q3: E frozen in mu-metal cells */
                if (k == 'n') {
                    orc_nb nb = nb_forward(L, iBc, j);
                    curl_forward(2, &nb, H_t, NULL, l, W);
                    if (j[2] == iI) { /* F:293-296 */
                        const double *src = Hi + 2 * (j[0] + L0 * j[1]);
                        W[0] += sgn * src[1];
                        W[1] -= sgn * src[0];
                    }
                    for (int c = 0; c < 3; c++)
                        E_t[c + 3 * l] = Ae[c + 3 * l] * E_t[c + 3 * l] + Be[c + 3 * l] * W[c];
                    curl_forward(2, &nb, Hs, addr, l, W);
                    for (int c = 0; c < 3; c++) E_t[c + 3 * l] += Be[c + 3 * l] * W[c];
                } else if (k == 'a') { /* F:300-307 */
                    orc_nb nb = nb_forward(L, iBc, j);
                    for (int half = 0; half < 2; half++) {
                        const int64_t s = half == 0 ? 3 * l : 3 * n + 3 * (int64_t)addr[l];
                        curl_forward(half, &nb, H_t, NULL, l, W);
                        for (int c = 0; c < 3; c++)
                            E_t[s + c] = Ae[s + c] * E_t[s + c] + Be[s + c] * W[c];
                        curl_forward(half, &nb, Hs, addr, l, W);
                        for (int c = 0; c < 3; c++) E_t[s + c] += Be[s + c] * W[c];
                    }
                } else { /* 'e' or 'b': Q1, the metal E update takes the BACKWARD curl (F:310) */
                    orc_nb nb = nb_backward(L, iBc, j);
                    const int64_t s = 3 * n + 3 * (int64_t)addr[l];
                    curl_backward(2, &nb, H_t, NULL, l, W);
                    for (int c = 0; c < 3; c++) E_t[c + 3 * l] += 0.5 * E_t[s + c];
                    for (int c = 0; c < 3; c++)
                        E_t[s + c] = Ae[s + c] * E_t[s + c] + Be[s + c] * W[c];
                    for (int c = 0; c < 3; c++) E_t[c + 3 * l] += 0.5 * E_t[s + c];
                }
            }
    }

    /* Running single-frequency DFT on the two monitor planes, main slots (F:318-328). */
    /* A plane index of -1 switches that monitor off (an extension of this build, used
     * where PML on the x/y faces puts PML cells in every plane). */
    for (int m = 0; m < 2; m++)
        for (int64_t q = 0; q < L0 * L1 && iBTIin[m] >= 0; q++)
            if (Cont[q + L0 * L1 * iBTIin[m]] == 'a') return 1 + m;
    const double e0r = ExpN2[0], e0i = ExpN2[1], e1r = ExpN2[2], e1i = ExpN2[3];
    for (int m = 0; m < 2; m++) {
        if (iBTIin[m] < 0) continue;
        double *HM = m == 0 ? HB : HT, *EM = m == 0 ? EB : ET;
        for (int64_t q = 0; q < L0 * L1; q++)
            for (int v = 0; v < 2; v++) {
                const int64_t c = v + 3 * (q + L0 * L1 * iBTIin[m]);
                const int64_t o = 2 * (v + 2 * q);
                HM[o] += e0r * H_t[c];
                HM[o + 1] += e0i * H_t[c];
                EM[o] += e1r * E_t[c];
                EM[o + 1] += e1i * E_t[c];
            }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------- */
/* Time constants, phase table and envelope (F:93-99, F:452-475, F:515-524)            */
/* ---------------------------------------------------------------------------------- */

/*
 * Smallest N with omega*Dt_min*N > 2*pi (long-double compare), then the time constants
 * that make the discrete d/dt and time average exact at omega (F:515-524).
 * Returns N; writes iR = 1/tau_R, iI = 1/tau_I and the step Dt.
 */
int orc_set_tau(double omega, double min_Dt, double *invtau_R, double *invtau_I, double *Dt_out)
{
    int N = 0;
    while (omega * min_Dt * N <= ORC_2PI) N++;
    const double dphi = ORC_2PI / (double)N;
    const double Dt = dphi / omega;
    *invtau_R = (0.5 * dphi) / (Dt * sin(0.5 * dphi));
    *invtau_I = (0.5 * omega) / (cos(0.5 * dphi));
    if (Dt_out) *Dt_out = Dt;
    return N;
}

/* ExpN[2t] = e^{i dphi t}, ExpN[2t+1] = e^{i dphi (t + 1/2)}, t < N (F:93-99). */
void orc_expn_table(int N, double *ExpN /* 2N complex */)
{
    const double dphi = ORC_2PI / (double)N;
    for (int t = 0; t < N; t++) {
        double complex a = cexp(CMPLX(0., dphi * (double)t));
        double complex b = cexp(CMPLX(0., dphi * (0.5 + (double)t)));
        ExpN[4 * t + 0] = creal(a);
        ExpN[4 * t + 1] = cimag(a);
        ExpN[4 * t + 2] = creal(b);
        ExpN[4 * t + 3] = cimag(b);
    }
}

/*
 * Source envelope (F:452-475): sin^2 ramp-up over tp[0] periods, 1 for tp[1] periods,
 * cos^2 ramp-down over tp[0] periods, 0 until tp[2]; normalised by 1/((tp0+tp1)*N).
 * Returns NAN where the reference exits ("f_env error!").
 */
double orc_set_env(int N, const int tp[3], int j_half, int j_time, int j_period)
{
    const int R = tp[0], C = tp[1], S = tp[2];
    double f;
    if (j_period < R || (R + C <= j_period && j_period < 2 * R + C)) {
        const double dph = ORC_2PI / ((double)4 * R);
        const double dphs = ORC_2PI / ((double)4 * R * N);
        const double p = j_period < R ? (double)j_period : (double)j_period - (R + C);
        double complex slow = cexp(CMPLX(0., dph * p));
        double complex fast = cexp(CMPLX(0., dphs * (0.5 * (double)j_half + (double)j_time)));
        double complex prod = slow * fast;
        f = j_period < R ? cimag(prod) : creal(prod);
    } else if (R <= j_period && j_period < R + C) {
        f = 1.;
    } else if (2 * R + C <= j_period && j_period < S) {
        f = 0.;
    } else {
        return NAN;
    }
    const double norm = 1. / ((double)1 * (R + C) * N);
    return norm * f * f;
}

/* ---------------------------------------------------------------------------------- */
/* PML profile recovery and removal (F:822-926)                                        */
/* ---------------------------------------------------------------------------------- */

/* Complex permeability/permittivity arrays are interleaved double[2*3n]. */
static inline double complex cget(const double *a, int64_t i)
{
    return CMPLX(a[2 * i], a[2 * i + 1]);
}
static inline void cset(double *a, int64_t i, double complex z)
{
    a[2 * i] = creal(z);
    a[2 * i + 1] = cimag(z);
}

/*
 * sigma[off_d + 2*j + h] = tan(arg) of the stretch along axis d at integer (h=0) or
 * half (h=1) position j, sampled only on the centre lines of the grid (Q7), and only
 * where the three stretched components carry the same phase (F:858-926).
 * off = (0, 2*L0, 2*L0 + 2*L1).  Entries that fail the test are left untouched.
 */
void orc_set_sigma(const int L[4], const double *mu, const double *eps, double *sigma)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2];
#define IDX(c, a, b, z) ((c) + 3 * ((a) + L0 * (b) + L0 * L1 * (z)))
    /* A sample of axis d compares the phases of three stretched components: two that
     * carry +arg(1 + i*sigma_d) and one of the other family that carries -arg. */
    {
        const int64_t z = L2 / 2, y = L1 / 2;
        for (int64_t x = 0; x < L0; x++) {
            double a1 = carg(cget(mu, IDX(1, x, y, z)));
            if (a1 != 0.) {
                double a2 = carg(cget(mu, IDX(2, x, y, z)));
                double a0 = -carg(cget(eps, IDX(0, x, y, z)));
                if (fabs(a0 - a1) < ORC_EMIN && fabs(a1 - a2) < ORC_EMIN && fabs(a2 - a0) < ORC_EMIN)
                    sigma[0 + 2 * x] = tan(a0);
            }
            a1 = carg(cget(eps, IDX(1, x, y, z)));
            if (a1 != 0.) {
                double a2 = carg(cget(eps, IDX(2, x, y, z)));
                double a0 = -carg(cget(mu, IDX(0, x, y, z)));
                if (fabs(a0 - a1) < ORC_EMIN && fabs(a1 - a2) < ORC_EMIN && fabs(a2 - a0) < ORC_EMIN)
                    sigma[1 + 2 * x] = tan(a0);
            }
        }
    }
    {
        const int64_t z = L2 / 2, x = L0 / 2;
        for (int64_t y = 0; y < L1; y++) {
            double a2 = carg(cget(mu, IDX(2, x, y, z)));
            if (a2 != 0.) {
                double a0 = carg(cget(mu, IDX(0, x, y, z)));
                double a1 = -carg(cget(eps, IDX(1, x, y, z)));
                if (fabs(a0 - a1) < ORC_EMIN && fabs(a1 - a2) < ORC_EMIN && fabs(a2 - a0) < ORC_EMIN)
                    sigma[2 * L0 + 0 + 2 * y] = tan(a0);
            }
            a2 = carg(cget(eps, IDX(2, x, y, z)));
            if (a2 != 0.) {
                double a0 = carg(cget(eps, IDX(0, x, y, z)));
                double a1 = -carg(cget(mu, IDX(1, x, y, z)));
                if (fabs(a0 - a1) < ORC_EMIN && fabs(a1 - a2) < ORC_EMIN && fabs(a2 - a0) < ORC_EMIN)
                    sigma[2 * L0 + 1 + 2 * y] = tan(a0);
            }
        }
    }
    {
        const int64_t y = L1 / 2, x = L0 / 2;
        for (int64_t z = 0; z < L2; z++) {
            double a0 = carg(cget(mu, IDX(0, x, y, z)));
            if (a0 != 0.) {
                double a1 = carg(cget(mu, IDX(1, x, y, z)));
                double a2 = -carg(cget(eps, IDX(2, x, y, z)));
                if (fabs(a0 - a1) < ORC_EMIN && fabs(a1 - a2) < ORC_EMIN && fabs(a2 - a0) < ORC_EMIN)
                    sigma[2 * L0 + 2 * L1 + 0 + 2 * z] = tan(a0);
            }
            a0 = carg(cget(eps, IDX(0, x, y, z)));
            if (a0 != 0.) {
                double a1 = carg(cget(eps, IDX(1, x, y, z)));
                double a2 = -carg(cget(mu, IDX(2, x, y, z)));
                if (fabs(a0 - a1) < ORC_EMIN && fabs(a1 - a2) < ORC_EMIN && fabs(a2 - a0) < ORC_EMIN)
                    sigma[2 * L0 + 2 * L1 + 1 + 2 * z] = tan(a0);
            }
        }
    }
#undef IDX
}

/*
 * Divide the stretch s_{c+1} s_{c+2} / s_c (s = 1 + i*sigma) back out of mu (integer
 * samples over the two others, half sample of its own axis) and eps (the opposite),
 * and make the result real where the stretch was complex (F:822-853).
 * Returns 0, 1 (mu phase left, "mu error in remove_sigma") or 2 (eps phase left).
 */
int orc_remove_sigma(const int L[4], double *mu, double *eps, const double *sigma)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2];
    double complex s[6]; /* s[h + 2*d] = 1 + i*sigma_d at integer (h=0) / half (h=1) */
    for (int64_t z = 0; z < L2; z++) {
        for (int h = 0; h < 2; h++) s[h + 4] = 1. + I * sigma[2 * L0 + 2 * L1 + h + 2 * z];
        for (int64_t y = 0; y < L1; y++) {
            for (int h = 0; h < 2; h++) s[h + 2] = 1. + I * sigma[2 * L0 + h + 2 * y];
            for (int64_t x = 0; x < L0; x++) {
                for (int h = 0; h < 2; h++) s[h] = 1. + I * sigma[h + 2 * x];
                const int64_t base = 3 * (x + L0 * y + L0 * L1 * z);
                for (int fam = 0; fam < 2; fam++) {
                    /* fam 0: mu uses s_{c+1}(0) s_{c+2}(0) / s_c(1); fam 1: eps, halves swapped */
                    double *arr = fam == 0 ? mu : eps;
                    const int hi = fam == 0 ? 0 : 1, ho = 1 - hi;
                    double complex ratio[3];
                    for (int c = 0; c < 3; c++) {
                        const double complex sd = s[ho + 2 * c];
                        ratio[c] = s[hi + 2 * ((c + 1) % 3)] * s[hi + 2 * ((c + 2) % 3)] * conj(sd) /
                                   (creal(sd) * creal(sd) + cimag(sd) * cimag(sd));
                    }
                    for (int c = 0; c < 3; c++) {
                        const double complex r = ratio[c];
                        double complex v = cget(arr, base + c);
                        v *= conj(r) / (creal(r) * creal(r) + cimag(r) * cimag(r));
                        if (cimag(r) != 0.) {
                            if (fabs(carg(v)) > ORC_EMIN) return 1 + fam;
                            v = creal(v);
                        }
                        cset(arr, base + c, v);
                    }
                }
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------- */
/* Cell classes and split-slot addresses (F:188-233)                                   */
/* ---------------------------------------------------------------------------------- */

/*
 * Cont[l]: 'a' when any of the cell's six sigma samples exceeds 10*DBL_EPSILON; then,
 * per component in order, a metal test that later components may overwrite (Q5):
 * 'b' both tan(arg mu) and tan(arg eps) < -EMIN, else 'h' (mu) or 'e' (eps).
 * addr[l]: running index over non-'n' cells in l order; 'n' cells get Lnotn.
 * Returns Lnotn.
 */
int orc_set_address(const int L[4], const double *sigma, const double *mu, const double *eps,
                    char *Cont, int *addr)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2], n = L0 * L1 * L2;
    for (int64_t z = 0; z < L2; z++)
        for (int64_t y = 0; y < L1; y++)
            for (int64_t x = 0; x < L0; x++) {
                const double sg[6] = {sigma[2 * x], sigma[2 * x + 1], sigma[2 * L0 + 2 * y],
                                      sigma[2 * L0 + 2 * y + 1], sigma[2 * L0 + 2 * L1 + 2 * z],
                                      sigma[2 * L0 + 2 * L1 + 2 * z + 1]};
                const int64_t l = x + L0 * y + L0 * L1 * z;
                char k = 'n';
                for (int i = 0; i < 6; i++)
                    if (fabs(sg[i]) > ORC_EMIN) k = 'a';
                for (int c = 0; c < 3; c++) {
                    const int mu_bad = tan(carg(cget(mu, c + 3 * l))) < -ORC_EMIN;
                    const int eps_bad = tan(carg(cget(eps, c + 3 * l))) < -ORC_EMIN;
                    if (mu_bad && eps_bad) k = 'b';
                    else if (mu_bad) k = 'h';
                    else if (eps_bad) k = 'e';
                }
                Cont[l] = k;
            }
    int Lnotn = 0;
    for (int64_t l = 0; l < n; l++)
        if (Cont[l] != 'n') Lnotn++;
    int next = 0;
    for (int64_t l = 0; l < n; l++) addr[l] = Cont[l] != 'n' ? next++ : Lnotn;
    return Lnotn;
}

/* ---------------------------------------------------------------------------------- */
/* Update coefficients (F:106-143)                                                     */
/* ---------------------------------------------------------------------------------- */

/*
 * Main slots: B = 1/(iR*Re x + iI*(Re x*sig + Im x)), A = (iR*Re x - iI*(...))*B with
 * sig = sigma of axis c+1 (integer samples for H, half samples for E).  'a' split
 * slots: the same with axis c+2.  Metal aux slots (StableForm.pdf p.3):
 * B = 1/(Re x/2 - iR*(Im x/2)/iI), A = -(Re x/2 + iR*(Im x/2)/iI)*B.
 * Output arrays hold 3*(n + Lnotn) doubles and must be zeroed by the caller.
 */
void orc_build_coeffs(const int L[4], const double *mu, const double *eps, const double *sigma,
                      const char *Cont, const int *addr, double iR, double iI, double *Ah,
                      double *Bh, double *Ae, double *Be)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2], n = L0 * L1 * L2;
    for (int64_t z = 0; z < L2; z++)
        for (int64_t y = 0; y < L1; y++)
            for (int64_t x = 0; x < L0; x++) {
                const int64_t l = x + L0 * y + L0 * L1 * z;
                /* sg[h + 2*d]: sigma of axis d at this cell, integer (h=0) / half (h=1) */
                const double sg[6] = {sigma[2 * x], sigma[2 * x + 1], sigma[2 * L0 + 2 * y],
                                      sigma[2 * L0 + 2 * y + 1], sigma[2 * L0 + 2 * L1 + 2 * z],
                                      sigma[2 * L0 + 2 * L1 + 2 * z + 1]};
                for (int c = 0; c < 3; c++) {
                    const int64_t i = c + 3 * l;
                    const double mr = mu[2 * i], mi = mu[2 * i + 1];
                    const double er = eps[2 * i], ei = eps[2 * i + 1];
                    const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
                    Bh[i] = 1. / (iR * mr + iI * (mr * sg[0 + 2 * d1] + mi));
                    Ah[i] = (iR * mr - iI * (mr * sg[0 + 2 * d1] + mi)) * Bh[i];
                    Be[i] = 1. / (iR * er + iI * (er * sg[1 + 2 * d1] + ei));
                    Ae[i] = (iR * er - iI * (er * sg[1 + 2 * d1] + ei)) * Be[i];
                    const int64_t s = c + 3 * (int64_t)addr[l] + 3 * n;
                    if (Cont[l] == 'a') {
                        Bh[s] = 1. / (iR * mr + iI * (mr * sg[0 + 2 * d2] + mi));
                        Ah[s] = (iR * mr - iI * (mr * sg[0 + 2 * d2] + mi)) * Bh[s];
                        Be[s] = 1. / (iR * er + iI * (er * sg[1 + 2 * d2] + ei));
                        Ae[s] = (iR * er - iI * (er * sg[1 + 2 * d2] + ei)) * Be[s];
                    }
                    if (Cont[l] == 'e' || Cont[l] == 'b') {
                        Be[s] = 1. / (0.5 * er - iR * (0.5 * ei / iI));
                        Ae[s] = -(0.5 * er + iR * (0.5 * ei / iI)) * Be[s];
                    }
                    if (Cont[l] == 'h' || Cont[l] == 'b') {
                        Bh[s] = 1. / (0.5 * mr - iR * (0.5 * mi / iI));
                        Ah[s] = -(0.5 * mr + iR * (0.5 * mi / iI)) * Bh[s];
                    }
                }
            }
}

/* ---------------------------------------------------------------------------------- */
/* Source synthesis and the time loop (F:153-167)                                      */
/* ---------------------------------------------------------------------------------- */

/*
 * Per-step source planes: Ei = 2*env(1/2, t)*Re(conj(ExpN[2t+1])*Einc),
 * Hi = 2*env(0, t+1)*Re(conj(ExpN[2((t+1)%N)])*Hinc) (F:155-158; Q9: t+1 unwrapped in
 * the envelope, wrapped in the phase).  Also returns the two envelope values.
 */
void orc_source_planes(int64_t nplane, int N, const int tp[3], int t, int P, const double *ExpN,
                       const double *Hinc, const double *Einc, double *Hi, double *Ei,
                       double *envH, double *envE)
{
    const double fE = orc_set_env(N, tp, 1, t, P);
    const double fH = orc_set_env(N, tp, 0, t + 1, P);
    const double complex eE = CMPLX(ExpN[4 * t + 2], ExpN[4 * t + 3]);
    const int tn = (t + 1) % N;
    const double complex eH = CMPLX(ExpN[4 * tn], ExpN[4 * tn + 1]);
    for (int64_t k = 0; k < nplane; k++) {
        Ei[k] = (2.) * fE * creal(conj(eE) * CMPLX(Einc[2 * k], Einc[2 * k + 1]));
        Hi[k] = (2.) * fH * creal(conj(eH) * CMPLX(Hinc[2 * k], Hinc[2 * k + 1]));
    }
    if (envH) *envH = fH;
    if (envE) *envE = fE;
}

/*
 * The time loop of F:153-167 on caller-owned arrays: fields must be zeroed and sized
 * 3*(n+Lnotn+1); monitors zeroed, 2*L0*L1 complex each.  Runs periods [P0, P1) of
 * the envelope (the full run is P0 = 0, P1 = tp[2]).  sums[2] accumulates sum00 and
 * sum05 (F:159-164).  Returns the first nonzero orc_main_calc status, else 0.
 */
int orc_run(const int L[4], const double *Ah, const double *Bh, const double *Ae,
            const double *Be, const int iBc[3], double *H_t, double *E_t, const int iBTIin[4],
            const double *Hinc, const double *Einc, int N, const int tp[3], int P0, int P1,
            double *HB, double *EB, double *HT, double *ET, const char *Cont, const int *addr,
            double sums[2])
{
    const int64_t np = 2 * (int64_t)L[0] * L[1];
    double *ExpN = malloc(sizeof(double) * 4 * N);
    double *Hi = malloc(sizeof(double) * np);
    double *Ei = malloc(sizeof(double) * np);
    int rc = 0;
    orc_expn_table(N, ExpN);
    for (int P = P0; P < P1 && rc == 0; P++)
        for (int t = 0; t < N && rc == 0; t++) {
            double fH, fE;
            orc_source_planes(np, N, tp, t, P, ExpN, Hinc, Einc, Hi, Ei, &fH, &fE);
            if (sums) {
                sums[0] += fH;
                sums[1] += fE;
            }
            rc = orc_main_calc(L, Ah, Bh, Ae, Be, iBc, H_t, E_t, iBTIin, Hi, Ei,
                               &ExpN[4 * ((t + 1) % N)], HB, EB, HT, ET, Cont, addr);
        }
    free(ExpN);
    free(Hi);
    free(Ei);
    return rc;
}

/*
 * Point current source on E component c (0, 1, 2 = x, y, z) of cell l, a class-'n' cell:
 * after the whole step (both sweeps and the monitors), E_c(l) += Be_c(l) * s, i.e. s is
 * added to the curl of that step's E update, E <- A E + B (curl H + s), rounded as
 * (A E + B curl H) + B s.
 */
void orc_point_add(double *E_t, const double *Be, int c, int64_t l, double s)
{
    E_t[c + 3 * l] += Be[c + 3 * l] * s;
}

/*
 * orc_run with a point source as well: after step k (k = P*N + t, counted from step 0 of
 * the envelope), E_c(l) += Be_c(l) * amp[k] for k < namp.  The sheet source and the
 * monitors run as in orc_run (zero Hinc, Einc for a point source alone).
 */
int orc_run_point(const int L[4], const double *Ah, const double *Bh, const double *Ae,
                  const double *Be, const int iBc[3], double *H_t, double *E_t, const int iBTIin[4],
                  const double *Hinc, const double *Einc, int N, const int tp[3], int P0, int P1,
                  double *HB, double *EB, double *HT, double *ET, const char *Cont, const int *addr,
                  double sums[2], int pc, int64_t pl, const double *amp, int64_t namp)
{
    const int64_t np = 2 * (int64_t)L[0] * L[1];
    double *ExpN = malloc(sizeof(double) * 4 * N);
    double *Hi = malloc(sizeof(double) * np);
    double *Ei = malloc(sizeof(double) * np);
    int rc = 0;
    orc_expn_table(N, ExpN);
    for (int P = P0; P < P1 && rc == 0; P++)
        for (int t = 0; t < N && rc == 0; t++) {
            double fH, fE;
            orc_source_planes(np, N, tp, t, P, ExpN, Hinc, Einc, Hi, Ei, &fH, &fE);
            if (sums) {
                sums[0] += fH;
                sums[1] += fE;
            }
            rc = orc_main_calc(L, Ah, Bh, Ae, Be, iBc, H_t, E_t, iBTIin, Hi, Ei,
                               &ExpN[4 * ((t + 1) % N)], HB, EB, HT, ET, Cont, addr);
            const int64_t k = (int64_t)P * N + t;
            if (rc == 0 && k < namp) orc_point_add(E_t, Be, pc, pl, amp[k]);
        }
    free(ExpN);
    free(Hi);
    free(Ei);
    return rc;
}

/* ---------------------------------------------------------------------------------- */
/* Mode projection (S_calc, F:437-445)                                                 */
/* ---------------------------------------------------------------------------------- */

/*
 * SB[col] = sum_r conj(tX[r,col])*HBx_r + conj(tX[r+P,col])*HBy_r
 *         + conj(tX[r+NA,col])*(-EBy_r) + conj(tX[r+P+NA,col])*EBx_r,  P = L0*L1,
 * for the ncol requested columns cols[]; tilXi columns are given densely as
 * tcols[k*2NA*2 ...] (2NA complex rows each, interleaved).
 */
void orc_s_project(const int L[4], int ncol, const double *tcols, const double *HB,
                   const double *EB, double *SB)
{
    const int64_t P = (int64_t)L[0] * L[1], NA = 2 * P;
    for (int k = 0; k < ncol; k++) {
        const double *col = tcols + (int64_t)k * 2 * NA * 2;
        double complex acc = 0.;
        for (int64_t r = 0; r < P; r++) {
            acc += conj(cget(col, r)) * cget(HB, 0 + 2 * r);
            acc += conj(cget(col, r + P)) * cget(HB, 1 + 2 * r);
            acc += conj(cget(col, r + NA)) * (-cget(EB, 1 + 2 * r));
            acc += conj(cget(col, r + P + NA)) * cget(EB, 0 + 2 * r);
        }
        SB[2 * k] = creal(acc);
        SB[2 * k + 1] = cimag(acc);
    }
}

/* ---------------------------------------------------------------------------------- */
/* Discrete adjoint of the step (no reference counterpart; checks the GPU adjoint)     */
/* ---------------------------------------------------------------------------------- */

/*
 * Transposed curl pieces: for every entry of P that curl_*(half, P, map) reads with
 * coefficient kappa into W[c], add kappa*u[c] to lP at that entry.
 */
static inline double *lslot(double *P, const int *map, int c, int64_t q)
{
    return map ? &P[c + 3 * (int64_t)map[q]] : &P[c + 3 * q];
}

static void curl_backward_T(int half, const orc_nb *nb, double *lP, const int *map, int64_t l, const double u[3])
{
    for (int c = 0; c < 3; c++) {
        int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        if (half != 1) {
            *lslot(lP, map, c2, l) += u[c];
            *lslot(lP, map, c2, l + nb->off[c1]) += -nb->w[c1] * u[c];
        }
        if (half != 0) {
            *lslot(lP, map, c1, l) += -u[c];
            *lslot(lP, map, c1, l + nb->off[c2]) += nb->w[c2] * u[c];
        }
    }
}

static void curl_forward_T(int half, const orc_nb *nb, double *lP, const int *map, int64_t l, const double u[3])
{
    for (int c = 0; c < 3; c++) {
        int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        if (half != 1) {
            *lslot(lP, map, c2, l + nb->off[c1]) += nb->w[c1] * u[c];
            *lslot(lP, map, c2, l) += -u[c];
        }
        if (half != 0) {
            *lslot(lP, map, c1, l + nb->off[c2]) += -nb->w[c2] * u[c];
            *lslot(lP, map, c1, l) += u[c];
        }
    }
}

/*
 * Reverse of one orc_main_calc step for grids of 'n' and 'a' cells: on entry lH, lE
 * hold dJ/d(state after the step), on exit dJ/d(state before it).  The step is
 * affine in the fields (the source does not depend on them), so this is its
 * transpose.  The zero slot of lH/lE is cleared.  Returns 1 if a metal cell is
 * present (not supported), else 0.
 */
int orc_step_adjoint(const int L[4], const double *Ah, const double *Bh, const double *Ae, const double *Be,
                     const int iBc[3], double *lH, double *lE, const char *Cont, const int *addr, int Lnotn)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2], n = L0 * L1 * L2;
    double *lHs = lH + 3 * n, *lEs = lE + 3 * n;
    int j[3];
    for (int64_t l = 0; l < n; l++)
        if (Cont[l] != 'n' && Cont[l] != 'a') return 1;
    /* reverse of the E sweep: E_new = Ae*E_old + Be*(curl_f(H) + curl_f_map(Hs) [+ src]) */
    for (j[2] = 0; j[2] < L2; j[2]++)
        for (j[1] = 0; j[1] < L1; j[1]++)
            for (j[0] = 0; j[0] < L0; j[0]++) {
                const int64_t l = j[0] + L0 * j[1] + L0 * L1 * j[2];
                orc_nb nb = nb_forward(L, iBc, j);
                if (Cont[l] == 'n') {
                    double u[3];
                    for (int c = 0; c < 3; c++) u[c] = Be[c + 3 * l] * lE[c + 3 * l];
                    curl_forward_T(2, &nb, lH, NULL, l, u);
                    curl_forward_T(2, &nb, lHs, addr, l, u);
                    for (int c = 0; c < 3; c++) lE[c + 3 * l] *= Ae[c + 3 * l];
                } else {
                    for (int half = 0; half < 2; half++) {
                        const int64_t s = half == 0 ? 3 * l : 3 * n + 3 * (int64_t)addr[l];
                        double u[3];
                        for (int c = 0; c < 3; c++) u[c] = Be[s + c] * lE[s + c];
                        curl_forward_T(half, &nb, lH, NULL, l, u);
                        curl_forward_T(half, &nb, lHs, addr, l, u);
                        for (int c = 0; c < 3; c++) lE[s + c] *= Ae[s + c];
                    }
                }
            }
    for (int c = 0; c < 3; c++) lHs[3 * (int64_t)Lnotn + c] = 0.;
    /* reverse of the H sweep: H_new = Ah*H_old - Bh*(curl_b(E) + curl_b_map(Es) [+ src]) */
    for (j[2] = 0; j[2] < L2; j[2]++)
        for (j[1] = 0; j[1] < L1; j[1]++)
            for (j[0] = 0; j[0] < L0; j[0]++) {
                const int64_t l = j[0] + L0 * j[1] + L0 * L1 * j[2];
                orc_nb nb = nb_backward(L, iBc, j);
                if (Cont[l] == 'n') {
                    double u[3];
                    for (int c = 0; c < 3; c++) u[c] = -Bh[c + 3 * l] * lH[c + 3 * l];
                    curl_backward_T(2, &nb, lE, NULL, l, u);
                    curl_backward_T(2, &nb, lEs, addr, l, u);
                    for (int c = 0; c < 3; c++) lH[c + 3 * l] *= Ah[c + 3 * l];
                } else {
                    for (int half = 0; half < 2; half++) {
                        const int64_t s = half == 0 ? 3 * l : 3 * n + 3 * (int64_t)addr[l];
                        double u[3];
                        for (int c = 0; c < 3; c++) u[c] = -Bh[s + c] * lH[s + c];
                        curl_backward_T(half, &nb, lE, NULL, l, u);
                        curl_backward_T(half, &nb, lEs, addr, l, u);
                        for (int c = 0; c < 3; c++) lH[s + c] *= Ah[s + c];
                    }
                }
            }
    for (int c = 0; c < 3; c++) lEs[3 * (int64_t)Lnotn + c] = 0.;
    return 0;
}

/*
 * Objective and gradient of a whole run (FP64, all states kept; small grids only).
 * Forward: the loop of F:153-167 from zero fields for periods [0, tp[2]) -- K steps.
 * Objective: J = sum_{v = x, y} sum_q |D_v(q)|^2, D_v(q) = sum_k e^{i*w*(k+1.5)}
 * E_v(q, zT) after step k (the E monitor's phase convention, F:323-325); E_v is the
 * main slot, and main + split (the total field) in PML cells.
 * Gradient: gradc[3n] = dJ/d(eps_c of each cell), through the E update coefficients
 * of 'n' cells (dBe/deps, dAe/deps from the coefficient formula, F:120-121); other
 * cells get 0.  eps: complex[3n] as given to orc_build_coeffs; sigma likewise.
 * Returns J, or NAN for unsupported input (metal cells).
 */
double orc_adjoint_run(const int L[4], const double *Ah, const double *Bh, const double *Ae, const double *Be,
                       const int iBc[3], const int iBTIin[4], const double *Hinc, const double *Einc, int N,
                       const int tp[3], const char *Cont, const int *addr, int Lnotn, const double *eps,
                       const double *sigma, double iR, double iI, int zT, double w, double *gradc)
{
    const int64_t L0 = L[0], L1 = L[1], n = L0 * L1 * L[2], plane = L0 * L1;
    const int64_t flen = 3 * (n + Lnotn + 1);
    const long K = (long)tp[2] * N;
    for (int64_t l = 0; l < n; l++)
        if (Cont[l] != 'n' && Cont[l] != 'a') return NAN;
    double *states = calloc((size_t)(K + 1) * 2 * flen, sizeof(double)); /* (H, E) after each step */
    double *ExpN = malloc(sizeof(double) * 4 * N);
    double *Hi = malloc(sizeof(double) * 2 * plane), *Ei = malloc(sizeof(double) * 2 * plane);
    double *mons = calloc(16 * plane, sizeof(double));
    orc_expn_table(N, ExpN);
    for (long k = 0; k < K; k++) {
        const int t = (int)(k % N), P = (int)(k / N);
        double *H = states + (k + 1) * 2 * flen, *E = H + flen;
        memcpy(H, states + k * 2 * flen, sizeof(double) * 2 * flen);
        orc_source_planes(2 * plane, N, tp, t, P, ExpN, Hinc, Einc, Hi, Ei, NULL, NULL);
        const int bti[4] = {-1, -1, iBTIin[2], iBTIin[3]};
        orc_main_calc(L, Ah, Bh, Ae, Be, iBc, H, E, bti, Hi, Ei, &ExpN[4 * ((t + 1) % N)], mons, mons + 4 * plane,
                      mons + 8 * plane, mons + 12 * plane, Cont, addr);
    }
    /* D and J */
    double *D = calloc(4 * plane, sizeof(double)); /* [v][q] complex */
    for (long k = 0; k < K; k++) {
        const double *E = states + (k + 1) * 2 * flen + flen;
        const double th = w * ((double)k + 1.5), cs = cos(th), sn = sin(th);
        for (int v = 0; v < 2; v++)
            for (int64_t q = 0; q < plane; q++) {
                const int64_t l = q + plane * zT;
                double x = E[v + 3 * l]; /* main slot; the total field in PML cells */
                if (Cont[l] == 'a') x += E[3 * n + v + 3 * (int64_t)addr[l]];
                D[2 * (v * plane + q)] += cs * x;
                D[2 * (v * plane + q) + 1] += sn * x;
            }
    }
    double J = 0.;
    for (int64_t i = 0; i < 2 * plane; i++) J += D[2 * i] * D[2 * i] + D[2 * i + 1] * D[2 * i + 1];
    /* reverse sweep */
    double *lH = calloc(flen, sizeof(double)), *lE = calloc(flen, sizeof(double));
    memset(gradc, 0, sizeof(double) * 3 * n);
    const int64_t off[3] = {0, 2 * L0, 2 * L0 + 2 * L1};
    for (long k = K - 1; k >= 0; k--) {
        /* objective sample of the state after step k */
        const double th = w * ((double)k + 1.5), cs = cos(th), sn = sin(th);
        for (int v = 0; v < 2; v++)
            for (int64_t q = 0; q < plane; q++) {
                const int64_t l = q + plane * zT;
                const double g = 2. * (D[2 * (v * plane + q)] * cs + D[2 * (v * plane + q) + 1] * sn);
                lE[v + 3 * l] += g;
                if (Cont[l] == 'a') lE[3 * n + v + 3 * (int64_t)addr[l]] += g;
            }
        /* gradient of step k: E_new = Ae*E_old + Be*W in 'n' cells, W = (E_new - Ae*E_old)/Be */
        const double *Eo = states + k * 2 * flen + flen, *En = states + (k + 1) * 2 * flen + flen;
        for (int64_t l = 0; l < n; l++) {
            if (Cont[l] != 'n') continue;
            const int64_t pos[3] = {l % L0, (l / L0) % L1, l / plane};
            for (int c = 0; c < 3; c++) {
                const int64_t i = c + 3 * l;
                const int d = (c + 1) % 3;
                const double s = sigma[off[d] + 1 + 2 * pos[d]];
                const double er = eps[2 * i], ei = eps[2 * i + 1];
                const double den = iR * er + iI * (er * s + ei);
                const double B = Be[i], A = Ae[i];
                const double dB = -(iR + iI * s) * B * B;
                const double dA = (iR - iI * s) * B + (iR * er - iI * (er * s + ei)) * dB;
                (void)den;
                const double W = (En[i] - A * Eo[i]) / B;
                gradc[i] += lE[i] * (dA * Eo[i] + dB * W);
            }
        }
        orc_step_adjoint(L, Ah, Bh, Ae, Be, iBc, lH, lE, Cont, addr, Lnotn);
    }
    free(states);
    free(ExpN);
    free(Hi);
    free(Ei);
    free(mons);
    free(D);
    free(lH);
    free(lE);
    return J;
}
