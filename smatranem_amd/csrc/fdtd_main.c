/*
 * fdtd_main.c -- the drop-in FDTD executable: `FDTD <input.txt>`.
 *
 * Replaces the reference's FDTD program (FDTDcalc/FDTDsplitfieldMetal.c, main at
 * F:32-187): reads the same input.txt keys and Data/ files, prints the same stderr
 * report, and exits with EXIT_FAILURE on the same conditions with the same messages.
 * The time loop runs on the GPU through libsmx_fdtd.so (include/smx_fdtd.h).
 *
 * Differences from the reference, by design:
 *  - Mode files: only the 2*N_prop columns that are printed are read, streamed block by
 *    block; the dense 2NA x 2NA tilXi / Xi matrices (F:435, F:490) are never formed.
 *  - check_Xi (F:495-496, a BLAS residual printed only when > 1e-15) is not computed.
 *  - SMX_FDTD_PRECISION=f32 in the environment selects the FP32 plan (default f64,
 *    bit-identical to the reference).
 *  - SMX_FDTD_PROJECT=gpu projects the monitors on the mode columns on the device
 *    (smx_plan_project) instead of on the host in the reference's order; the |S|^2
 *    printed then differ from the reference's by rounding only.
 *  - SMX_FDTD_SETUP_DUMP=<file> is a dry run that needs no GPU: after the plane indices
 *    are reported, the program writes what it read from the files to <file> and exits
 *    with status 0 (layout at dump_setup).
 * The file formats are the reference's (and its writers', Medium3DSi/Med3DSi.c:210-258,
 * PMLset/PMLcalc.c:331, ModeCalc/Mcalc.c:191-228), so the scan formats below follow
 * them.
 */
#include <complex.h>
#include <ctype.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "smx_fdtd.h"

#define LINE 1024                                                      /* constant.h:6 */
#define TWO_PI_LD (2.L * 3.1415926535897932384626433832795028841971L) /* constant.h:15 */

static void die(const char *msg)
{
    fputs(msg, stderr);
    exit(EXIT_FAILURE);
}

/* ---------------------------------------------------------------------------------- */
/* input.txt (F:759-817)                                                              */
/* ---------------------------------------------------------------------------------- */

typedef struct {
    char prefix[LINE];
    char plane_b[LINE], plane_t[LINE], plane_inc[LINE];
    int bc[3];
    int tp[3];
    int inc_mode;
    double dt_factor;
} run_input;

static void strip_leading_space(char *s)
{
    size_t k = 0;
    while (s[k] && isspace((unsigned char)s[k])) k++;
    if (k) memmove(s, s + k, strlen(s + k) + 1);
}

static void cut_at_comma(char *s)
{
    char *c = strchr(s, ',');
    if (c) *c = '\0';
}

static void read_input(const char *path, run_input *in)
{
    FILE *f = fopen(path, "r");
    if (!f) die("open error!: open input-file0!\n");
    char buf[LINE];
    int found = 0;
    while (found < 4 && fgets(buf, sizeof buf, f)) {
        strip_leading_space(buf);
        if (!strncmp(buf, "Prefix", 6) && sscanf(buf, "%*[^=] %*[=] %s", in->prefix) == 1) {
            cut_at_comma(in->prefix);
            found++;
        } else if (!strncmp(buf, "BoundaryCondition", 17) &&
                   sscanf(buf, "%*[^=] %*[=] %d %*[^=] %*[=] %d %*[^=] %*[=] %d", &in->bc[0], &in->bc[1],
                          &in->bc[2]) == 3) {
            found++;
        } else if (!strncmp(buf, "ModeP", 5) &&
                   sscanf(buf, "%*[^=] %*[=] %*[^=] %*[=] %*[^=] %*[=] %s", in->plane_b) == 1) {
            cut_at_comma(in->plane_b);
            if (sscanf(buf, "%*[^=] %*[=] %*[^=] %*[=] %*[^=] %*[=] %*[^=] %*[=] %s", in->plane_t) != 1)
                sscanf(in->plane_b, "%s", in->plane_t);
            else
                cut_at_comma(in->plane_t);
            found++;
        } else if (!strncmp(buf, "FDTD", 4) &&
                   sscanf(buf, "%*[^=] %*[=] %d %*[^=] %*[=] %d %*[^=] %*[=] %d %*[^=] %*[=] %s %*[^=] %*[=] %d "
                               "%*[^=] %*[=] %lf",
                          &in->tp[0], &in->tp[1], &in->tp[2], in->plane_inc, &in->inc_mode, &in->dt_factor) == 6) {
            cut_at_comma(in->plane_inc);
            found++;
        }
    }
    if (fclose(f) != 0) die("fclose error after input_file0!\n");
    if (found != 4) die("4 control commands cannot be read in input_file0!");
}

/* Plane names B<k>, T<k>, <k> to the global z index (F:703-710). */
static int plane_index(const int L[4], const char *s)
{
    int k = -1;
    if (s[0] == 'B') sscanf(s + 1, "%d", &k);
    else if (s[0] == 'T') {
        sscanf(s + 1, "%d", &k);
        k += L[2] - L[3];
    } else {
        sscanf(s, "%d", &k);
        k += L[3];
    }
    return k;
}

/* ---------------------------------------------------------------------------------- */
/* Med files (F:665-752)                                                              */
/* ---------------------------------------------------------------------------------- */

static void slice_file(const char *prefix, const int L[4], int z, char *name)
{
    if (z < L[3]) snprintf(name, LINE, "%s_MedB%d.dat", prefix, z);
    else if (z < L[2] - L[3]) snprintf(name, LINE, "%s_Med%d.dat", prefix, z - L[3]);
    else snprintf(name, LINE, "%s_MedT%d.dat", prefix, z - (L[2] - L[3]));
}

/* Grid sizes from <prefix>_Med0.dat and omega = 2*pi/(wavelength*k0), which the
 * bMode file must repeat exactly (F:724-752). */
static void read_sizes(const char *prefix, int L[4], double *omega)
{
    char name[LINE], buf[LINE];
    snprintf(name, sizeof name, "%s_Med0.dat", prefix);
    FILE *f = fopen(name, "r");
    if (!f) {
        fprintf(stderr, "open error for %s\n", name);
        exit(EXIT_FAILURE);
    }
    int need = 3;
    double wavelength = 0., k0 = 0.;
    while (need > 0 && fgets(buf, sizeof buf, f)) {
        if (sscanf(buf, "%*[^<] %*[<] %d %*[^<] %*[<] %d %*[^<] %*[<] %d %*[^<] %*[<] %*[^=] %*[=] %d", &L[0], &L[1],
                   &L[2], &L[3]) == 4)
            need--;
        else if (!strncmp(buf, "# wavelength", 12) && sscanf(buf, "%*[^=] %*[=] %lf", &wavelength) == 1)
            need--;
        else if (!strncmp(buf, "# k_0", 5) && sscanf(buf, "%*[^=] %*[=] %*[^=] %*[=] %lf", &k0) == 1)
            need--;
    }
    if (fclose(f) != 0) die("fclose error after reading the Med0 file!\n");
    if (need > 0) die("L,M,N,N_outer can not be read at Scatterer @ input_file0!\n");
    fprintf(stderr, "wavelength = %.5e, k_0 = %.5e\n", wavelength, k0);
    *omega = TWO_PI_LD / (wavelength * k0);

    snprintf(name, sizeof name, "%s_bMode.dat", prefix);
    f = fopen(name, "r");
    if (!f) {
        fprintf(stderr, "open error for %s\n", name);
        exit(EXIT_FAILURE);
    }
    double om = 0.;
    int got = 0;
    while (!got && fgets(buf, sizeof buf, f))
        if (!strncmp(buf, "# Omega", 7) && sscanf(buf, "%*[^=] %*[=] %lf", &om) == 1) got = 1;
    if (fclose(f) != 0) die("fclose error after reading the bMode file!\n");
    if (!got) die("Omega can not be read in input_Lf!\n");
    if (*omega != om) {
        fprintf(stderr, "Omega = %.5e != %.5e in input_Lf!\n", *omega, om);
        exit(EXIT_FAILURE);
    }
}

/* One slice: 2*L0*L1 data lines, j0 fastest then j1, j2 in {0, 1} innermost;
 * j2 = 0 carries mu_x, mu_y, eps_z and j2 = 1 eps_y, eps_x, mu_z, each (Re + i Im)*f. */
static double read_slice(const char *name, const int L[4], double complex *mu, double complex *eps)
{
    FILE *f = fopen(name, "r");
    if (!f) {
        fprintf(stderr, "open error for %s\n", name);
        exit(EXIT_FAILURE);
    }
    char buf[LINE];
    int have_dt = 0, x = 0, y = 0, h = 0, lines = 0;
    double min_sqdt = 0.;
    while (fgets(buf, sizeof buf, f)) {
        if (!strncmp(buf, "# l <", 5) &&
            sscanf(buf, "%*[^<] %*[<] %*[^<] %*[<] %*[^<] %*[<] %*[^<] %*[<] %*[^=] %*[=] %*[^=] %*[=] %*[^=] %*[=] %lf",
                   &min_sqdt) == 1)
            have_dt = 1;
        int a, b, c;
        double r0, i0, r1, i1, r2, i2, f0, f1, f2;
        if (sscanf(buf, "%d %d %d %lf %lf %lf %lf %lf %lf %lf %lf %lf", &a, &b, &c, &r0, &i0, &r1, &i1, &r2, &i2, &f0,
                   &f1, &f2) != 12)
            continue;
        if (a != x || b != y || c != h) {
            fputs("Data cannot be read in input_data2!\n", stderr);
            fprintf(stderr, "%s\n", buf);
            exit(EXIT_FAILURE);
        }
        const int64_t q = 3 * ((int64_t)x + (int64_t)L[0] * y);
        if (h == 0) {
            mu[q + 0] = (r0 + I * i0) * f0;
            mu[q + 1] = (r1 + I * i1) * f1;
            eps[q + 2] = (r2 + I * i2) * f2;
        } else {
            eps[q + 1] = (r0 + I * i0) * f0;
            eps[q + 0] = (r1 + I * i1) * f1;
            mu[q + 2] = (r2 + I * i2) * f2;
        }
        if (++h == 2) {
            h = 0;
            if (++x == L[0]) {
                x = 0;
                y++;
            }
        }
        lines++;
    }
    if (fclose(f) != 0) die("fclose error in input_data!\n");
    if (!have_dt) die("min_sqDt can not be read in input_data!\n");
    if (lines != 2 * L[0] * L[1]) {
        fprintf(stderr, "j_count2 = %d != %d in input_data!\n", lines, 2 * L[0] * L[1]);
        exit(EXIT_FAILURE);
    }
    return min_sqdt;
}

/* ---------------------------------------------------------------------------------- */
/* Mode files (F:479-661), streamed                                                   */
/* ---------------------------------------------------------------------------------- */

/* Number of propagating ("real") modes from a Mode file (F:632-661). */
static int real_mode_count(const char *name, const int L[4])
{
    FILE *f = fopen(name, "r");
    if (!f) {
        fprintf(stderr, "open error!: open %s!\n", name);
        exit(EXIT_FAILURE);
    }
    char buf[LINE];
    int headers = 0, n_real = 0, n_cplx = 0, thetas = 0;
    while (fgets(buf, sizeof buf, f)) {
        int a, b;
        double ra, ia, rb, ib;
        if (!strncmp(buf, "# Eigen-value", 13) && sscanf(buf, "%*[^=] %*[=] %d %*[^=] %*[=] %d", &n_real, &n_cplx) == 2)
            headers++;
        else if (!strncmp(buf, "theta[", 6) &&
                 sscanf(buf, "%*[^[] %*[[] %d %*[^(] %*[(] %lf %*[,] %lf %*[^[] %*[[] %d %*[^(] %*[(] %lf %*[,] %lf", &a,
                        &ra, &ia, &b, &rb, &ib) == 6) {
            if (a != thetas) die("theta cannot be read in get_RealNum05!\n");
            thetas++;
        }
    }
    if (headers != 1 || n_real + n_cplx != 2 * L[0] * L[1] || thetas != 2 * L[0] * L[1])
        die("RealNum05 cannot be read in get_RealNum05!\n");
    if (fclose(f) != 0) die("fclose error in get_RealNum05!\n");
    return n_real;
}

/* Stream the first `nblocks` column blocks of an Xi file (F:557-628): block j holds
 * rows r < 2NA of columns j and j+NA.  cb(j, row, a, b) receives them in file order. */
typedef void (*xi_cb)(void *ctx, int j, int row, double complex a, double complex b);

static void stream_xi(const char *name, int NA, int nblocks, xi_cb cb, void *ctx)
{
    FILE *f = fopen(name, "r");
    if (!f) {
        fprintf(stderr, "open error!: open %s!\n", name);
        exit(EXIT_FAILURE);
    }
    char buf[LINE];
    for (int k = 0; k < 2; k++)
        if (!fgets(buf, sizeof buf, f)) die("set_tilPhi error!\n");
    for (int j = 0; j < nblocks; j++) {
        if (!fgets(buf, sizeof buf, f)) die("set_tilPhi error!\n");
        int idx = -1;
        if (sscanf(buf, "%*[^[] %*[[] %d", &idx) == 1 && idx != j) {
            fprintf(stderr, "dummyi = %d, jcolumn = %d!\n", idx, j);
            exit(EXIT_FAILURE);
        }
        for (int r = 0; r < 2 * NA; r++) {
            double a0, a1, b0, b1;
            if (!fgets(buf, sizeof buf, f)) die("set_tilPhi error!\n");
            if (sscanf(buf, "%lf %lf %lf %lf", &a0, &a1, &b0, &b1) != 4) {
                fprintf(stderr, "From %s, data cannot be read at column = %d, row = %d!\n", name, j, r);
                fprintf(stderr, "%s\n", buf);
                exit(EXIT_FAILURE);
            }
            cb(ctx, j, r, a0 + I * a1, b0 + I * b1);
        }
    }
    fclose(f);
}

typedef struct {
    int column;  /* wanted column of Xi */
    int NA;
    double complex *col;
} col_ctx;

static void grab_column(void *p, int j, int row, double complex a, double complex b)
{
    col_ctx *c = p;
    if (j == c->column % c->NA) c->col[row] = c->column < c->NA ? a : b;
}

/* set_incident (F:479-511): the incident column of the nearer port's Xi file. */
static void incident_profile(const char *prefix, const int L[4], int iB, int iT, int iI, int mode,
                             double complex *Hinc, double complex *Einc)
{
    char xi[LINE], md[LINE];
    const char *side = abs(iB - iI) < abs(iT - iI) ? "b" : "t";
    snprintf(xi, sizeof xi, "%s_%sXi.dat", prefix, side);
    snprintf(md, sizeof md, "%s_%sMode.dat", prefix, side);
    const int P = L[0] * L[1], NA = 2 * P;
    const int n_real = real_mode_count(md, L);
    if (mode > n_real) {
        fprintf(stderr, "inci_mode = %d > %d\n", mode, n_real);
        exit(EXIT_FAILURE);
    }
    if (mode == 0) {
        fprintf(stderr, "inci_mode = %d!\n", mode);
        exit(EXIT_FAILURE);
    }
    const int column = mode > 0 ? mode - 1 : -mode - 1 + NA;
    double complex *col = calloc(2 * (size_t)NA, sizeof *col);
    col_ctx ctx = {column, NA, col};
    stream_xi(xi, NA, column % NA + 1, grab_column, &ctx);
    for (int r = NA; r < 2 * NA; r++)
        if (cimag(col[r]) != 0.) die("incident wave is not real number!\n");
    for (int r = 0; r < P; r++) {
        Hinc[0 + 2 * r] = col[r];
        Hinc[1 + 2 * r] = col[r + P];
        Einc[1 + 2 * r] = -col[r + 2 * P];
        Einc[0 + 2 * r] = col[r + 3 * P];
    }
    free(col);
}

typedef struct {
    int NA, n_real;
    const double complex *HB, *EB;
    double complex *S; /* [2*n_real]: S[m], then S[m + NA] at index n_real + m */
    double complex *colA, *colB;
} proj_ctx;

/* Block j < n_real of the Xi file gives tilXi columns j and j+NA (F:617-619): rows
 * rotated by NA, column j+NA negated.  Project once a block is complete (F:437-445). */
static void project_block(void *p, int j, int row, double complex a, double complex b)
{
    proj_ctx *c = p;
    const int NA = c->NA, P = NA / 2;
    const int r = (row + NA) % (2 * NA);
    c->colA[r] = a;
    c->colB[r] = -b;
    if (row != 2 * NA - 1 || j >= c->n_real) return;
    for (int which = 0; which < 2; which++) {
        const double complex *t = which ? c->colB : c->colA;
        double complex acc = 0.;
        for (int q = 0; q < P; q++) {
            acc += conj(t[q]) * c->HB[0 + 2 * q];
            acc += conj(t[q + P]) * c->HB[1 + 2 * q];
            acc += conj(t[q + NA]) * (-c->EB[1 + 2 * q]);
            acc += conj(t[q + P + NA]) * c->EB[0 + 2 * q];
        }
        c->S[which * c->n_real + j] = acc;
    }
}

/* S_calc (F:419-448) for the printed columns; returns N_prop, fills S. */
static int project(const char *prefix, const int L[4], int plane, const double complex *HB,
                   const double complex *EB, double complex **S_out)
{
    char xi[LINE], md[LINE];
    if (plane >= 0 && plane < L[3]) {
        snprintf(xi, sizeof xi, "%s_bXi.dat", prefix);
        snprintf(md, sizeof md, "%s_bMode.dat", prefix);
    } else if (plane >= L[2] - L[3] && plane < L[2]) {
        snprintf(xi, sizeof xi, "%s_tXi.dat", prefix);
        snprintf(md, sizeof md, "%s_tMode.dat", prefix);
    } else {
        fprintf(stderr, "iB or iT = %d \n", plane);
        exit(EXIT_FAILURE);
    }
    const int NA = 2 * L[0] * L[1];
    const int n_real = real_mode_count(md, L);
    proj_ctx c = {NA, n_real, HB, EB, calloc(2 * (size_t)(n_real + 1), sizeof(double complex)),
                  calloc(2 * (size_t)NA, sizeof(double complex)), calloc(2 * (size_t)NA, sizeof(double complex))};
    stream_xi(xi, NA, n_real, project_block, &c);
    free(c.colA);
    free(c.colB);
    *S_out = c.S;
    return n_real;
}

typedef struct {
    int NA, n_real;
    double complex *modes; /* [2*n_real][2NA]: tilXi columns j < n_real, then j + NA */
} cols_ctx;

/* Block j < n_real gives tilXi columns j and j+NA (F:617-619): rows rotated by NA, the
 * second column negated (as project_block forms them). */
static void collect_block(void *p, int j, int row, double complex a, double complex b)
{
    cols_ctx *c = p;
    if (j >= c->n_real) return;
    const int r = (row + c->NA) % (2 * c->NA);
    c->modes[(size_t)j * 2 * c->NA + r] = a;
    c->modes[(size_t)(c->n_real + j) * 2 * c->NA + r] = -b;
}

static void check(int rc);

/* S_calc (F:419-448) for the printed columns of one port, projected on the device from
 * the plan's monitors (smx_plan_project); port 0 bottom, 1 top. */
static int project_gpu(smx_plan *plan, const char *prefix, const int L[4], int plane, int port,
                       double complex **S_out)
{
    char xi[LINE], md[LINE];
    if (plane >= 0 && plane < L[3]) {
        snprintf(xi, sizeof xi, "%s_bXi.dat", prefix);
        snprintf(md, sizeof md, "%s_bMode.dat", prefix);
    } else if (plane >= L[2] - L[3] && plane < L[2]) {
        snprintf(xi, sizeof xi, "%s_tXi.dat", prefix);
        snprintf(md, sizeof md, "%s_tMode.dat", prefix);
    } else {
        fprintf(stderr, "iB or iT = %d \n", plane);
        exit(EXIT_FAILURE);
    }
    const int NA = 2 * L[0] * L[1];
    const int n_real = real_mode_count(md, L);
    double complex *S = calloc(2 * (size_t)(n_real + 1), sizeof *S);
    if (n_real > 0) {
        cols_ctx c = {NA, n_real, calloc((size_t)2 * n_real * 2 * (size_t)NA, sizeof(double complex))};
        double complex *smon = calloc(2 * (size_t)2 * n_real, sizeof *smon);
        if (!c.modes || !smon) die("out of memory in S_calc\n");
        stream_xi(xi, NA, n_real, collect_block, &c);
        check(smx_plan_project(plan, 2 * n_real, (const double *)c.modes, (double *)smon, NULL));
        memcpy(S, smon + (size_t)port * 2 * n_real, sizeof *S * 2 * (size_t)n_real);
        free(c.modes);
        free(smon);
    }
    *S_out = S;
    return n_real;
}

static void check(int rc)
{
    if (rc == SMX_OK) return;
    if (rc == SMX_ERR_MONITOR_PML || rc == SMX_ERR_ENVELOPE) {
        fprintf(stderr, "%s\n", smx_last_error());
        exit(EXIT_FAILURE);
    }
    fprintf(stderr, "FDTD: %s: %s\n", smx_status_string(rc), smx_last_error());
    exit(EXIT_FAILURE);
}

/* The dry run's file: int32 L[4], BoundaryCondition[3], periods[3], incident mode, iB, iT,
 * iI; float64 Dt_factor, omega, min_Dt; then mu and eps as read from the Med files
 * (complex128[3n] each, fx applied, before set_sigma), all little-endian. */
static void dump_setup(const char *path, const int L[4], const run_input *in, const int planes[3], double omega,
                       double min_dt, const double complex *mu, const double complex *eps)
{
    FILE *f = fopen(path, "wb");
    if (!f) die("open error for SMX_FDTD_SETUP_DUMP\n");
    const int ints[14] = {L[0],     L[1],     L[2],     L[3],        in->bc[0], in->bc[1], in->bc[2],
                          in->tp[0], in->tp[1], in->tp[2], in->inc_mode, planes[0], planes[1], planes[2]};
    const double dbl[3] = {in->dt_factor, omega, min_dt};
    const size_t n3 = 3 * (size_t)L[0] * L[1] * L[2];
    if (fwrite(ints, sizeof ints, 1, f) != 1 || fwrite(dbl, sizeof dbl, 1, f) != 1 ||
        fwrite(mu, sizeof *mu, n3, f) != n3 || fwrite(eps, sizeof *eps, n3, f) != n3 || fclose(f) != 0)
        die("write error for SMX_FDTD_SETUP_DUMP\n");
}

int main(int argc, char **argv)
{
    time_t now = time(0);
    fprintf(stderr, "# Start time of FDTD_SpiltField = %s\n", ctime(&now));
    if (argc != 2) die("error: number of files \n");
    if (!strncmp(argv[1], "-v", 2) || !strcmp(argv[1], "--version")) {
        fprintf(stderr, "The '%s' creates scattering matrix column vector by FDTD on an AMD GPU.\n", argv[0]);
        fprintf(stderr, "smatranem_amd FDTD (MI355X / gfx950), drop-in for FDTDcalc/FDTDsplitfieldMetal.c "
                        "version 19.10.29.\n");
        exit(EXIT_SUCCESS);
    }

    run_input in;
    memset(&in, 0, sizeof in);
    in.dt_factor = 1.;
    read_input(argv[1], &in);
    int L[4];
    double omega = 0.;
    read_sizes(in.prefix, L, &omega);
    fprintf(stderr, "Ramp-period = %d, cw-period = %d, stop-period = %d\n", in.tp[0], in.tp[1], in.tp[2]);
    fprintf(stderr, "Dt_factor = %.5e, split field impl. does not use stable_factor.\n", in.dt_factor);

    const int64_t P = (int64_t)L[0] * L[1], n = P * L[2];
    double complex *mu = calloc(3 * (size_t)n, sizeof *mu), *eps = calloc(3 * (size_t)n, sizeof *eps);
    double min_dt = 0.;
    for (int z = 0; z < L[2]; z++) {
        char name[LINE];
        slice_file(in.prefix, L, z, name);
        const double sq = read_slice(name, L, mu + 3 * P * z, eps + 3 * P * z);
        if (z == 0 || min_dt > sqrt(sq)) min_dt = sqrt(sq);
    }

    const char *dump = getenv("SMX_FDTD_SETUP_DUMP");
    double complex *mu_read = NULL, *eps_read = NULL;
    if (dump && *dump) {
        mu_read = malloc(3 * (size_t)n * sizeof *mu_read);
        eps_read = malloc(3 * (size_t)n * sizeof *eps_read);
        if (!mu_read || !eps_read) die("out of memory for SMX_FDTD_SETUP_DUMP\n");
        memcpy(mu_read, mu, 3 * (size_t)n * sizeof *mu);
        memcpy(eps_read, eps, 3 * (size_t)n * sizeof *eps);
    }

    const int nsig = 2 * (L[0] + L[1] + L[2]);
    double *sigma = calloc((size_t)nsig, sizeof *sigma);
    smx_set_sigma(L, (const double *)mu, (const double *)eps, sigma);
    { /* the reference reports each recovered sample (F:868-920) */
        const int off[3] = {0, 2 * L[0], 2 * L[0] + 2 * L[1]};
        for (int d = 0; d < 3; d++)
            for (int j = 0; j < L[d]; j++)
                for (int h = 0; h < 2; h++)
                    if (sigma[off[d] + h + 2 * j] != 0.)
                        fprintf(stderr, "l%d = %d.%d, sigma = %.5e\n", d, j, h ? 5 : 0, sigma[off[d] + h + 2 * j]);
    }
    const int rs = smx_remove_sigma(L, (double *)mu, (double *)eps, sigma);
    if (rs == 1) die("mu error in remove_sigma\n");
    if (rs == 2) die("epsi error in remove_sigma\n");

    int iB = plane_index(L, in.plane_b), iT = plane_index(L, in.plane_t), iI = plane_index(L, in.plane_inc);
    if (iB == iT) {
        iT = (L[2] - 1) - iB;
        fputs("Change iT because iT = iB\n", stderr);
    }
    if (iB >= iT) {
        fprintf(stderr, "Target_B = %d for %s is larger than or equal to Target_T = %d for %s\n", iB, in.plane_b, iT,
                in.plane_t);
        exit(EXIT_FAILURE);
    }
    fprintf(stderr, "Start = %d for %s, Incident point = %d for %s, Incident mode = %+d, Target_T = %d for %s\n", iB,
            in.plane_b, iI, in.plane_inc, in.inc_mode, iT, in.plane_t);
    if (mu_read) {
        const int planes[3] = {iB, iT, iI};
        dump_setup(dump, L, &in, planes, omega, min_dt, mu_read, eps_read);
        return 0;
    }

    double complex *Hinc = calloc(2 * (size_t)P, sizeof *Hinc), *Einc = calloc(2 * (size_t)P, sizeof *Einc);
    incident_profile(in.prefix, L, iB, iT, iI, in.inc_mode, Hinc, Einc);

    double iR, iIv, Dt;
    const int N = smx_set_tau(omega, min_dt / in.dt_factor, &iR, &iIv, &Dt);
    fprintf(stderr, "omega = %.5e, min_Dt = %.5e, N_2pi = %d, Dt = %.5e, tau_R = %.5e, tau_I = %.5e\n", omega,
            min_dt / in.dt_factor, N, Dt, 1. / iR, 1. / iIv);

    char *Cont = calloc((size_t)n, 1);
    int *addr = calloc((size_t)n, sizeof *addr);
    const int Lnotn = smx_set_address(L, sigma, (const double *)mu, (const double *)eps, Cont, addr);
    fprintf(stderr, "L3d = %d, Lnotn = %d\n", (int)n, Lnotn);

    const char *prec = getenv("SMX_FDTD_PRECISION");
    smx_grid g;
    memset(&g, 0, sizeof g);
    memcpy(g.L, L, sizeof g.L);
    memcpy(g.iBc, in.bc, sizeof g.iBc);
    g.iBTIin[0] = iB;
    g.iBTIin[1] = iT;
    g.iBTIin[2] = iI;
    g.iBTIin[3] = in.inc_mode;
    g.precision = (prec && !strcmp(prec, "f32")) ? SMX_F32 : SMX_F64;
    g.Lnotn = Lnotn;
    g.invtau_R = iR;
    g.invtau_I = iIv;
    smx_plan *plan = NULL;
    check(smx_plan_create(&g, (const double *)mu, (const double *)eps, sigma, Cont, addr, &plan));
    free(mu);
    free(eps);
    free(sigma);
    free(Cont);
    free(addr);
    check(smx_plan_set_source(plan, (const double *)Hinc, (const double *)Einc, N, in.tp));
    double sums[2] = {0., 0.};
    check(smx_plan_run(plan, 0, in.tp[2], sums));
    fprintf(stderr, "sum00 = %.5e, sum05 = %.5e\n", sums[0], sums[1]);
    fprintf(stderr, "Bc[0] = %d, Bc[1] = %d, Bc[2] = %d\n", in.bc[0], in.bc[1], in.bc[2]);

    double complex *mon[4];
    for (int k = 0; k < 4; k++) mon[k] = calloc(2 * (size_t)P, sizeof(double complex));
    check(smx_plan_read_monitors(plan, (double *)mon[0], (double *)mon[1], (double *)mon[2], (double *)mon[3]));
    double complex *SB = NULL, *ST = NULL;
    int nB, nT;
    const char *proj = getenv("SMX_FDTD_PROJECT");
    if (proj && !strcmp(proj, "gpu")) { /* on the device, from the plan's monitors */
        nB = project_gpu(plan, in.prefix, L, iB, 0, &SB);
        nT = project_gpu(plan, in.prefix, L, iT, 1, &ST);
    } else { /* on the host, in the reference's summation order */
        nB = project(in.prefix, L, iB, mon[0], mon[1], &SB);
        nT = project(in.prefix, L, iT, mon[2], mon[3], &ST);
    }
    smx_plan_destroy(plan);
    free(Hinc);
    free(Einc);

    fprintf(stderr, "Bottom |S_(m,%+d)|^2 :\n", in.inc_mode);
    for (int m = 0; m < nB; m++)
        fprintf(stderr, "%+d, %.20e, %+d, %.20e\n", m + 1, creal(conj(SB[m]) * SB[m]), -(m + 1),
                creal(conj(SB[nB + m]) * SB[nB + m]));
    fprintf(stderr, "Top |S_(m,%+d)|^2 :\n", in.inc_mode);
    for (int m = 0; m < nT; m++)
        fprintf(stderr, "%+d, %.20e, %+d, %.20e\n", m + 1, creal(conj(ST[m]) * ST[m]), -(m + 1),
                creal(conj(ST[nT + m]) * ST[nT + m]));
    for (int k = 0; k < 4; k++) free(mon[k]);
    free(SB);
    free(ST);
    return 0;
}
