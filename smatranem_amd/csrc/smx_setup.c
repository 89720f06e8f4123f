/*
 * smx_setup.c -- host-side setup of the FDTD plan: time constants, envelope, phase
 * table, PML profile recovery, cell classes and update coefficients.
 *
 * These run once per run (or once per step for the envelope scalars) on the host, in
 * FP64, and must reproduce the reference's values bit for bit, because the FP64 plan
 * is checked bitwise against main_calc.  The file is C11 and is compiled with
 * gcc -O2 -std=c11 -ffp-contract=off, the reference's own compiler and flags
 * (FDTDcalc/comp.sh:3), so that each expression below rounds exactly as the
 * reference's does.  Expressions keep the reference's operation order on purpose;
 * the loops around them are this library's own.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "smx_fdtd.h"
#include "smx_setup.h"

#define SMX_TOL (10. * DBL_EPSILON)                                    /* constant.h:7 */
#define SMX_TWO_PI (2.L * 3.1415926535897932384626433832795028841971L) /* constant.h:15 */

/* set_tau, F:515-524. */
int smx_set_tau(double omega, double min_Dt, double *invtau_R, double *invtau_I, double *Dt)
{
    /* smallest N with omega*min_Dt*N > 2*pi, compared in long double (F:518) */
    int N = 0;
    while (omega * min_Dt * N <= SMX_TWO_PI) N++;
    const double phase_step = SMX_TWO_PI / (double)N;
    const double dt = phase_step / omega;
    *invtau_R = (0.5 * phase_step) / (dt * sin(0.5 * phase_step));
    *invtau_I = (0.5 * omega) / (cos(0.5 * phase_step));
    if (Dt) *Dt = dt;
    return N;
}

/* ExpN table, F:93-99: entry t holds e^{i*dphi*t} and e^{i*dphi*(t+1/2)}. */
void smx_expn_table(int N, double *ExpN)
{
    const double phase_step = SMX_TWO_PI / (double)N;
    for (int t = 0; t < N; t++) {
        const double complex whole = cexp(CMPLX(0., phase_step * (double)t));
        const double complex half = cexp(CMPLX(0., phase_step * (0.5 + (double)t)));
        double *e = ExpN + 4 * t;
        e[0] = creal(whole);
        e[1] = cimag(whole);
        e[2] = creal(half);
        e[3] = cimag(half);
    }
}

/* set_env, F:452-475: the branches are tested in the reference's order. */
double smx_set_env(int N, const int tp[3], int j_half, int j_time, int j_period)
{
    const int ramp = tp[0], cw = tp[1], stop = tp[2];
    double f;
    int ramping = 0, rising = 0;
    if (j_period < ramp) {
        ramping = rising = 1;
    } else if (j_period < ramp + cw) {
        f = 1.;
    } else if (j_period < 2 * ramp + cw) {
        ramping = 1;
    } else if (j_period < stop) {
        f = 0.;
    } else {
        return NAN;
    }
    if (ramping) {
        const double coarse = SMX_TWO_PI / ((double)4 * ramp);
        const double fine = SMX_TWO_PI / ((double)4 * ramp * N);
        const double p = rising ? (double)j_period : (double)j_period - (ramp + cw);
        const double complex ep = cexp(CMPLX(0., coarse * p));
        const double complex es = cexp(CMPLX(0., fine * (0.5 * (double)j_half + (double)j_time)));
        const double complex z = ep * es;
        f = rising ? cimag(z) : creal(z);
    }
    const double inv_norm = 1. / ((double)1 * (ramp + cw) * N);
    return inv_norm * f * f;
}

/* ------------------------------------------------------------------------------------
 * set_sigma, F:858-926.
 *
 * Each of the six profile samples (axis d, integer/half position h) is read on one
 * centre line of the grid from three stretched components: two that carry +arg of
 * the stretch and one that carries -arg.  The table names them; the sample is
 * tan(arg) where the three phases agree within 10*DBL_EPSILON and the first is nonzero.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    int lead_eps;  /* family of the first (+arg) component: 0 mu, 1 eps */
    int lead_c;    /* its component */
    int second_c;  /* second +arg component (same family as the lead) */
    int opp_c;     /* -arg component, in the other family */
} smx_sigma_probe;

/* index [d][h] */
static const smx_sigma_probe smx_probes[3][2] = {
    {{0, 1, 2, 0}, {1, 1, 2, 0}}, /* x: mu_y, mu_z vs eps_x;  eps_y, eps_z vs mu_x */
    {{0, 2, 0, 1}, {1, 2, 0, 1}}, /* y: mu_z, mu_x vs eps_y;  eps_z, eps_x vs mu_y */
    {{0, 0, 1, 2}, {1, 0, 1, 2}}, /* z: mu_x, mu_y vs eps_z;  eps_x, eps_y vs mu_z */
};

static double complex cat(const double *a, int64_t i) { return CMPLX(a[2 * i], a[2 * i + 1]); }

void smx_set_sigma(const int L[4], const double *mu, const double *eps, double *sigma)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2];
    const int64_t centre[3] = {L0 / 2, L1 / 2, L2 / 2};
    const int64_t stride[3] = {1, L0, L0 * L1};
    const int64_t offset[3] = {0, 2 * L0, 2 * L0 + 2 * L1};
    for (int d = 0; d < 3; d++) {
        int64_t base = 0;
        for (int e = 0; e < 3; e++)
            if (e != d) base += centre[e] * stride[e];
        for (int64_t j = 0; j < L[d]; j++) {
            const int64_t cell = base + j * stride[d];
            for (int h = 0; h < 2; h++) {
                const smx_sigma_probe *p = &smx_probes[d][h];
                const double *lead = p->lead_eps ? eps : mu;
                const double *opp = p->lead_eps ? mu : eps;
                const double a_lead = carg(cat(lead, p->lead_c + 3 * cell));
                if (a_lead == 0.) continue;
                const double a_second = carg(cat(lead, p->second_c + 3 * cell));
                const double a_opp = -carg(cat(opp, p->opp_c + 3 * cell));
                /* Arrange the three phases by the component they belong to (x, y, z);
                 * all three pairs must agree, and the sample is tan of the x one. */
                double ax, ay, az;
                if (d == 0) { ax = a_opp; ay = a_lead; az = a_second; }
                else if (d == 1) { ax = a_second; ay = a_opp; az = a_lead; }
                else { ax = a_lead; ay = a_second; az = a_opp; }
                if (fabs(ax - ay) < SMX_TOL && fabs(ay - az) < SMX_TOL && fabs(az - ax) < SMX_TOL)
                    sigma[offset[d] + h + 2 * j] = tan(ax);
            }
        }
    }
}

/* remove_sigma, F:822-853. */
int smx_remove_sigma(const int L[4], double *mu, double *eps, const double *sigma)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2];
    const int64_t off[3] = {0, 2 * L0, 2 * L0 + 2 * L1};
    for (int64_t z = 0; z < L2; z++)
        for (int64_t y = 0; y < L1; y++)
            for (int64_t x = 0; x < L0; x++) {
                const int64_t pos[3] = {x, y, z};
                double complex st[3][2]; /* stretch 1 + i*sigma, [axis][integer/half] */
                for (int d = 0; d < 3; d++)
                    for (int h = 0; h < 2; h++) st[d][h] = 1. + I * sigma[off[d] + h + 2 * pos[d]];
                const int64_t cell = 3 * (x + L0 * y + L0 * L1 * z);
                for (int fam = 0; fam < 2; fam++) {
                    double *arr = fam ? eps : mu;
                    const int h_pair = fam, h_self = 1 - fam; /* mu: s(0)s(0)/s(1); eps: s(1)s(1)/s(0) */
                    for (int c = 0; c < 3; c++) {
                        const double complex den = st[c][h_self];
                        const double complex ratio = st[(c + 1) % 3][h_pair] * st[(c + 2) % 3][h_pair] *
                                                     conj(den) /
                                                     (creal(den) * creal(den) + cimag(den) * cimag(den));
                        double complex v = cat(arr, cell + c);
                        v *= conj(ratio) / (creal(ratio) * creal(ratio) + cimag(ratio) * cimag(ratio));
                        if (cimag(ratio) != 0.) {
                            if (fabs(carg(v)) > SMX_TOL) return 1 + fam;
                            v = creal(v);
                        }
                        arr[2 * (cell + c)] = creal(v);
                        arr[2 * (cell + c) + 1] = cimag(v);
                    }
                }
            }
    return 0;
}

/* set_address, F:188-233. */
int smx_set_address(const int L[4], const double *sigma, const double *mu, const double *eps,
                    char *Cont, int *address_notn)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2], n = L0 * L1 * L2;
    const int64_t off[3] = {0, 2 * L0, 2 * L0 + 2 * L1};
    /* per-axis "in PML" flags: either sample of the axis above the tolerance */
    const int64_t Lmax = L0 > L1 ? (L0 > L2 ? L0 : L2) : (L1 > L2 ? L1 : L2);
    char *pml = calloc(3 * Lmax, 1);
    for (int d = 0; d < 3; d++)
        for (int64_t j = 0; j < L[d]; j++)
            pml[d * Lmax + j] = fabs(sigma[off[d] + 2 * j]) > SMX_TOL || fabs(sigma[off[d] + 2 * j + 1]) > SMX_TOL;
    int count = 0;
    for (int64_t l = 0; l < n; l++) {
        const int64_t x = l % L0, y = (l / L0) % L1, z = l / (L0 * L1);
        char k = (pml[x] || pml[Lmax + y] || pml[2 * Lmax + z]) ? 'a' : 'n';
        for (int c = 0; c < 3; c++) { /* later components overwrite earlier ones (Q5) */
            const int m = tan(carg(cat(mu, c + 3 * l))) < -SMX_TOL;
            const int e = tan(carg(cat(eps, c + 3 * l))) < -SMX_TOL;
            if (m && e) k = 'b';
            else if (m) k = 'h';
            else if (e) k = 'e';
        }
        Cont[l] = k;
        count += k != 'n';
    }
    free(pml);
    int next = 0;
    for (int64_t l = 0; l < n; l++) address_notn[l] = Cont[l] == 'n' ? count : next++;
    return count;
}

/*
 * Coefficients of one cell (F:117-139): out[0..2] Ah, [3..5] Bh, [6..8] Ae, [9..11] Be
 * of the main slots; out[12..23] the same four for the split ('a') or stable-form aux
 * (metal) slot, written only for the classes that use them.  mu3, eps3: complex[3];
 * sg[h + 2*d]: sigma of axis d at the cell, integer (h=0) / half (h=1) sample.
 */
void smx_cell_coeffs(char k, const double *mu3, const double *eps3, const double sg[6], double iR,
                     double iI, double out[24])
{
    double *Ah = out, *Bh = out + 3, *Ae = out + 6, *Be = out + 9;
    double *Ahs = out + 12, *Bhs = out + 15, *Aes = out + 18, *Bes = out + 21;
    for (int c = 0; c < 3; c++) {
        const double mr = mu3[2 * c], mi = mu3[2 * c + 1], er = eps3[2 * c], ei = eps3[2 * c + 1];
        /* H takes integer samples, E half samples; main slot axis c+1, split axis c+2 */
        const int d_main = (c + 1) % 3, d_split = (c + 2) % 3;
        const double sh_m = sg[0 + 2 * d_main], se_m = sg[1 + 2 * d_main];
        Bh[c] = 1. / (iR * mr + iI * (mr * sh_m + mi));
        Ah[c] = (iR * mr - iI * (mr * sh_m + mi)) * Bh[c];
        Be[c] = 1. / (iR * er + iI * (er * se_m + ei));
        Ae[c] = (iR * er - iI * (er * se_m + ei)) * Be[c];
        if (k == 'a') {
            const double sh_s = sg[0 + 2 * d_split], se_s = sg[1 + 2 * d_split];
            Bhs[c] = 1. / (iR * mr + iI * (mr * sh_s + mi));
            Ahs[c] = (iR * mr - iI * (mr * sh_s + mi)) * Bhs[c];
            Bes[c] = 1. / (iR * er + iI * (er * se_s + ei));
            Aes[c] = (iR * er - iI * (er * se_s + ei)) * Bes[c];
        }
        if (k == 'e' || k == 'b') { /* stable form, check_files/StableForm.pdf p.3 */
            Bes[c] = 1. / (0.5 * er - iR * (0.5 * ei / iI));
            Aes[c] = -(0.5 * er + iR * (0.5 * ei / iI)) * Bes[c];
        }
        if (k == 'h' || k == 'b') {
            Bhs[c] = 1. / (0.5 * mr - iR * (0.5 * mi / iI));
            Ahs[c] = -(0.5 * mr + iR * (0.5 * mi / iI)) * Bhs[c];
        }
    }
}

/* sigma samples of cell (x, y, z): sg[h + 2*d] */
void smx_cell_sigma(const int L[4], const double *sigma, int64_t x, int64_t y, int64_t z, double sg[6])
{
    const int64_t o1 = 2 * (int64_t)L[0], o2 = 2 * (int64_t)L[0] + 2 * (int64_t)L[1];
    sg[0] = sigma[2 * x];
    sg[1] = sigma[2 * x + 1];
    sg[2] = sigma[o1 + 2 * y];
    sg[3] = sigma[o1 + 2 * y + 1];
    sg[4] = sigma[o2 + 2 * z];
    sg[5] = sigma[o2 + 2 * z + 1];
}

/* Coefficient build, F:106-143. */
void smx_build_coeffs(const int L[4], const double *mu, const double *eps, const double *sigma,
                      const char *Cont, const int *address_notn, int Lnotn, double iR, double iI,
                      double *Ah, double *Bh, double *Ae, double *Be)
{
    (void)Lnotn;
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2], n = L0 * L1 * L2;
    for (int64_t l = 0; l < n; l++) {
        double sg[6], w[24];
        smx_cell_sigma(L, sigma, l % L0, (l / L0) % L1, l / (L0 * L1), sg);
        const char k = Cont[l];
        smx_cell_coeffs(k, mu + 6 * l, eps + 6 * l, sg, iR, iI, w);
        const int64_t s = 3 * (int64_t)address_notn[l] + 3 * n;
        for (int c = 0; c < 3; c++) {
            Ah[3 * l + c] = w[c];
            Bh[3 * l + c] = w[3 + c];
            Ae[3 * l + c] = w[6 + c];
            Be[3 * l + c] = w[9 + c];
            if (k == 'a' || k == 'h' || k == 'b') {
                Ah[s + c] = w[12 + c];
                Bh[s + c] = w[15 + c];
            }
            if (k == 'a' || k == 'e' || k == 'b') {
                Ae[s + c] = w[18 + c];
                Be[s + c] = w[21 + c];
            }
        }
    }
}
