/*
 * smx_fdtd.h -- C ABI of the MI355X split-field FDTD library (libsmx_fdtd.so).
 *
 * Drop-in for the time-stepper of SMatrAnEM's FDTD executable,
 * FDTDcalc/FDTDsplitfieldMetal.c (cited as F:<line>).  Plain C types only: pointers,
 * sizes and int status codes; complex arrays are interleaved (re, im) doubles, the
 * memory layout of C99 `double complex`, so a caller holding `double complex *` passes
 * it cast to `double *`.  All pointers are host pointers unless a name says `_dev`.
 *
 * Storage contract shared with the reference (F:151-152, F:188-233):
 *   n = L0*L1*L2; cell l = j0 + L0*j1 + L0*L1*j2.
 *   Fields H_t, E_t: double[3*(n + Lnotn + 1)]; main slots X[c + 3l], then split/aux
 *   slots X[3n + c + 3*address_notn[l]]; index Lnotn is the never-written zero slot.
 *   Coefficients Ah, Bh, Ae, Be: double[3*(n + Lnotn)], same indexing.
 *   Cont[l] in {'n','a','e','h','b'}; address_notn[l] = running index of non-'n' cells,
 *   Lnotn for 'n' cells.
 */
#ifndef SMX_FDTD_H
#define SMX_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Codes >= 100 are the reference's own fatal conditions; the drop-in
 * executable prints the reference's message for them and exits with EXIT_FAILURE. */
enum {
    SMX_OK = 0,
    SMX_ERR_ARG = 1,           /* invalid argument or inconsistent sizes */
    SMX_ERR_HIP = 2,           /* HIP runtime error (see smx_last_error) */
    SMX_ERR_NO_DEVICE = 3,     /* no HIP device visible */
    SMX_ERR_OOM = 4,           /* device allocation failed */
    SMX_ERR_STATE = 5,         /* call out of order (e.g. run before set_source) */
    SMX_ERR_MONITOR_PML = 100, /* a monitor plane holds a PML cell: "iBTIin[m] error!" F:320 */
    SMX_ERR_ENVELOPE = 101     /* step beyond stop-period: "f_env error!" F:471 */
};

/* Human-readable text of a status code, and the last HIP error text of this thread. */
const char *smx_status_string(int status);
const char *smx_last_error(void);

/* ------------------------------------------------------------------------------------
 * (i) Parity entry point: exactly main_calc (F:5-12, F:242-249).
 *
 * One leapfrog step on host arrays with the reference's argument meanings and layouts:
 * H sweep, E sweep, then the running DFT on the monitor planes iBTIin[0..1].  Runs the
 * plain FP64 HIP kernels; the arrays are copied to the device and back on every call.
 * Hi, Ei: double[2*L0*L1] source planes, (x, y) interleaved.  ExpN2: complex[2].
 * HB, EB, HT, ET: complex[2*L0*L1], accumulated in place.
 * Returns SMX_OK, or SMX_ERR_MONITOR_PML where the reference exits (F:320).
 * Replaces: main_calc, FDTDcalc/FDTDsplitfieldMetal.c:242.
 * ---------------------------------------------------------------------------------- */
int smx_main_calc(const int L[4], const double *Ah, const double *Bh, const double *Ae,
                  const double *Be, const int iBc[3], double *H_t, double *E_t,
                  const int iBTIin[4], const double *Hi, const double *Ei, const double *ExpN2,
                  double *HB, double *EB, double *HT, double *ET, const char *Cont,
                  const int *address_notn);

/*
 * No counterpart in the reference: smx_main_calc for callers that cannot allocate
 * device memory (code that forbids hipMalloc/hipFree, such as tight serving loops or
 * harnesses that own all device memory).  Same arguments, semantics and results as
 * smx_main_calc (bit-identical), but the arrays are staged in a fixed device array of
 * SMX_NOALLOC_MAX_BYTES that the library holds on each device it runs on until the
 * process ends, instead of buffers allocated per call.  A grid whose staged arrays
 * (4 coefficient and 2 field arrays, Cont, address_notn, source planes and monitors,
 * each rounded up to 256 bytes) exceed it is refused with SMX_ERR_ARG before any device
 * work.  Calls on one device share that array; the library serialises them with a
 * per-device lock, and work the caller queued on other streams is not ordered against
 * them.
 */
#define SMX_NOALLOC_MAX_BYTES ((size_t)64 << 20)
int smx_main_calc_noalloc(const int L[4], const double *Ah, const double *Bh, const double *Ae,
                          const double *Be, const int iBc[3], double *H_t, double *E_t,
                          const int iBTIin[4], const double *Hi, const double *Ei,
                          const double *ExpN2, double *HB, double *EB, double *HT, double *ET,
                          const char *Cont, const int *address_notn);

/* ------------------------------------------------------------------------------------
 * (ii) Production plan: device-resident state, many steps per call.
 *
 * A plan owns the device copies of fields, coefficients and monitors of one grid.
 * Precision SMX_F64 runs the plain FP64 kernels on the reference storage contract and
 * is bit-identical to main_calc.  SMX_F32 runs the fused FP32 kernels on a packed
 * layout (total fields plus split slots, per-cell FP32 coefficient words, per-axis PML
 * tables) with FP64 monitor accumulators.
 * ---------------------------------------------------------------------------------- */
typedef struct smx_plan smx_plan;

enum { SMX_F64 = 0, SMX_F32 = 1 };

/* Grid and run description: the scalars main() holds before its time loop (F:57-104). */
typedef struct {
    int L[4];       /* L0, L1, L2 (= scatterer slices + 2*outer), outer (F:57) */
    int iBc[3];     /* BoundaryCondition B0, B1, B2 in {-1, 0, 1} (F:46) */
    int iBTIin[4];  /* iB, iT (monitor planes), iI (source plane), incident mode (F:84) */
    int precision;  /* SMX_F64 or SMX_F32 */
    int Lnotn;      /* number of non-'n' cells (F:102-104) */
    double invtau_R, invtau_I; /* time constants from set_tau (F:92) */
} smx_grid;

/*
 * Create a plan from the setup state main() builds before its loop:
 *   mu, eps: complex[3n] after remove_sigma (F:76), per component (x, y, z), scale
 *            factor ratios included;
 *   sigma:   double[2*(L0+L1+L2)] PML profiles from set_sigma (F:75);
 *   Cont, address_notn: cell classes (F:104).
 * The F64 plan builds the reference's Ah/Bh/Ae/Be (F:106-143) on the host and uploads
 * them; the F32 plan packs its own layout.  Replaces: the setup of main, F:100-152.
 */
int smx_plan_create(const smx_grid *grid, const double *mu, const double *eps,
                    const double *sigma, const char *Cont, const int *address_notn,
                    smx_plan **out);

/*
 * Create an F64 plan directly from the four coefficient arrays main_calc receives
 * (F:106-107), e.g. from a caller that builds them itself.  grid->precision must be
 * SMX_F64; invtau_R/I are unused.
 */
int smx_plan_create_coeffs(const smx_grid *grid, const double *Ah, const double *Bh,
                           const double *Ae, const double *Be, const char *Cont,
                           const int *address_notn, smx_plan **out);

int smx_plan_destroy(smx_plan *plan);

/* Size in doubles of one field array in the reference layout: 3*(n + Lnotn + 1). */
size_t smx_plan_field_len(const smx_plan *plan);

/* Copy fields in/out in the reference layout (double[3*(n+Lnotn+1)]).  For F32 plans
 * the main slot is reconstructed as total minus split, in FP64. */
int smx_plan_write_fields(smx_plan *plan, const double *H_t, const double *E_t);
int smx_plan_read_fields(const smx_plan *plan, double *H_t, double *E_t);

/* Monitors HB, EB, HT, ET: complex[2*L0*L1] each, read, written or zeroed. */
int smx_plan_read_monitors(const smx_plan *plan, double *HB, double *EB, double *HT, double *ET);
int smx_plan_write_monitors(smx_plan *plan, const double *HB, const double *EB,
                            const double *HT, const double *ET);

/* One step with explicit host source planes and phases (main_calc's inputs). */
int smx_plan_step(smx_plan *plan, const double *Hi, const double *Ei, const double *ExpN2);

/*
 * Modal source (F:479-511 output): Hinc, Einc complex[2*L0*L1], (x, y) interleaved;
 * N_2pi steps per period and the envelope periods (ramp, cw, stop) (F:60, F:92).
 */
int smx_plan_set_source(smx_plan *plan, const double *Hinc, const double *Einc, int N_2pi,
                        const int time_period[3]);

/*
 * Run envelope periods [P0, P1) of the loop at F:153-167, synthesising the source on
 * the device.  sums[2] (may be NULL) accumulates the envelope sums sum00, sum05
 * printed at F:168.  Returns SMX_ERR_ENVELOPE where the reference exits (F:471).
 */
int smx_plan_run(smx_plan *plan, int P0, int P1, double sums[2]);

/*
 * Run global steps [step0, step0 + nsteps) of the same loop, step k being period
 * k / N_2pi, phase step k % N_2pi.  smx_plan_run(P0, P1) is
 * smx_plan_run_steps(P0*N_2pi, (P1-P0)*N_2pi).
 */
int smx_plan_run_steps(smx_plan *plan, long long step0, long long nsteps, double sums[2]);

/*
 * Batch: global steps [step0, step0 + nsteps) of every plan in plans[0 .. nplans-1]
 * (independent runs, each with its own grid, source and monitors), with the same
 * results as smx_plan_run_steps on each plan in turn.  Each plan runs on its own stream,
 * so kernels of different plans run concurrently and small grids share the device; over
 * the CW periods one captured HIP graph of every plan's steps is replayed
 * (smx_plan_run_batch_mode, SMX_BATCH_GRAPH).  sums (may be NULL) is double[2*nplans],
 * the envelope sums of each plan.  On an error return the plans may have advanced by
 * different numbers of steps.  No counterpart in the reference (its runs are separate
 * processes, run_all.sh).
 */
int smx_plan_run_batch(smx_plan *const *plans, int nplans, long long step0, long long nsteps, double *sums);

/*
 * Batch with an explicit mode.  smx_plan_run_batch is mode SMX_BATCH_GRAPH.
 *   SMX_BATCH_STREAMS: every step (or two-step pass) of every plan launched from the
 *     host, step-major, each plan on its own stream.
 *   SMX_BATCH_GRAPH: as STREAMS until every plan has entered a stretch where the
 *     per-step source and monitor scalars repeat (the CW periods of the envelope); then
 *     one segment of every plan (G steps, G a multiple of its N_2pi and of 4) is
 *     captured from all plan streams into one HIP graph, and that graph is replayed for
 *     as long as the scalars of every plan's next segment equal the captured ones bit
 *     for bit; the remaining steps run as in STREAMS.  Plans with a DFT set (or a
 *     segment longer than 1024 steps, or fewer than two segments to run) run as in
 *     STREAMS.  Results are those of STREAMS exactly.
 *   SMX_BATCH_FUSED: every two-step pass of all plans in ONE launch (two_step_multi, a
 *     grid of every plan's tiles), one launch per pass; an odd last step runs as in
 *     STREAMS.  Each plan runs the split-slot two-step pass (SMX_F32_TWO variant 2)
 *     whatever kernel it had selected, so the results are bit for bit those of each
 *     plan run alone with that kernel, DFT accumulators included (each plan's ring
 *     slot of a pass goes up with its step scalars; its folds run between passes, where
 *     the plan alone folds).  Needs FP32 plans without lossy/metal cells, with a DFT
 *     chunk of 2 or more where a DFT is set, and under 4 GiB per field group; otherwise
 *     the batch runs as SMX_BATCH_GRAPH.
 * graph_launches (may be NULL) receives the number of graph launches (SMX_BATCH_FUSED:
 * of two_step_multi launches).  The batch allocates its graph or tables, stream and
 * events per call and frees them before returning.  No counterpart in the reference.
 */
enum { SMX_BATCH_STREAMS = 0, SMX_BATCH_GRAPH = 1, SMX_BATCH_FUSED = 2 };
int smx_plan_run_batch_mode(smx_plan *const *plans, int nplans, long long step0, long long nsteps,
                            double *sums, int mode, int *graph_launches);

/* Wait for all queued work of the plan. */
int smx_plan_sync(smx_plan *plan);

/* Queue the folds of the DFT and monitor samples the plan still holds in its ring
 * (smx_plan_set_dft) on the plan's stream, without waiting; a no-op without a DFT.
 * Reads of the DFT or monitors do this themselves; a caller that times a run up to an
 * event on the plan's stream calls it before recording the event.  Folding earlier than
 * the chunk boundary changes only how the samples are grouped in the FP64 sums.  No
 * counterpart in the reference. */
int smx_plan_flush(smx_plan *plan);

/*
 * Extension: running DFT of the tangential fields on up to 8 z-planes at up to 32
 * frequencies (the reference's monitors, F:318-328, sample one frequency on two
 * planes).  w_per_step[f] = omega_f*Dt.  Run step k (k counted from 0 by
 * smx_plan_run_steps, or by successive smx_plan_step calls) adds
 * e^{i*w*(k+1)}*(Hx, Hy) and e^{i*w*(k+1.5)}*(Ex, Ey) of the main slots (the total
 * field in PML cells); with w = 2*pi/N_2pi this is the reference's monitor.  Samples
 * are buffered and folded into FP64 accumulators every `chunk` (1..256) steps; FP32 plans
 * running two steps per pass take the samples of both steps inside the pass (chunk >= 2).
 * While the DFT is on, an FP32 plan also buffers the monitor planes (up to 8 planes in
 * all) and folds the reference monitors from those samples with the reference's phases
 * instead of accumulating them per step (steps run by smx_plan_step, with the caller's
 * ExpN2, still accumulate per step); smx_plan_read_monitors folds what is pending first.
 * The monitors then differ from per-step accumulation only in the order of the sums.
 * nplanes = 0 switches the DFT off.  No counterpart in the reference.
 */
int smx_plan_set_dft(smx_plan *plan, int nplanes, const int *planes, int nfreq, const double *w_per_step,
                     int chunk);
/* Accumulated DFT, complex[nplanes][4 (Hx, Hy, Ex, Ey)][nfreq][L0*L1]. */
int smx_plan_read_dft(smx_plan *plan, double *out);

/*
 * Extension: a point current source on E component comp (0, 1, 2 = x, y, z) of cell
 * (cell[0], cell[1], cell[2]), which must be a lossless class-'n' cell.  After run step k
 * (counted as for smx_plan_set_dft) and before its DFT sample, E_comp += B * amp[k] for
 * k < nsteps, B the cell's E coefficient (Be of the reference layout, 1/(iR eps)): amp[k]
 * is added to the curl of step k's E update, E <- A E + B (curl H + amp[k]), rounded as
 * (A E + B curl H) + B amp[k] in FP64.  The reference's source is the modal sheet
 * (F:261-264, F:293-296); this is the Gaussian-pulse point excitation of a broadband run,
 * and runs beside the sheet (zero Hinc, Einc for the point source alone).  An FP32 plan
 * with a point source takes one step per pass, and a batch runs it eagerly (no graph, no
 * fused launch); the adjoint refuses it.  nsteps = 0 removes it.  No counterpart in the
 * reference.
 */
int smx_plan_set_point_source(smx_plan *plan, int comp, const int cell[3], long long nsteps,
                              const double *amp);

/*
 * The mode projection of S_calc (F:437-445) on the device: for each column t of tilXi
 * given (modes: complex[nmodes][4*L0*L1], rows in tilXi's order, F:598-628), the sum
 * over q of conj(t[q]) Hx[q] + conj(t[q+P]) Hy[q] + conj(t[q+2P]) (-Ey[q]) +
 * conj(t[q+3P]) Ex[q], P = L0*L1, of
 *   S_mon: complex[2][nmodes] -- the monitors (HB, EB) then (HT, ET); a monitor that
 *          is off gives zeros;
 *   S_dft: complex[nplanes][nfreq][nmodes] -- every plane and frequency of the DFT set
 *          by smx_plan_set_dft (SMX_ERR_STATE without one).
 * Either output may be NULL, not both.  Pending samples are folded first.  FP64 with a
 * fixed summation order (rows in chunks of 256 reduced by a tree, chunks in order): it
 * differs from the reference's sequential sum only by rounding.  Replaces the loop of
 * S_calc, F:437-445, for device-resident monitors.
 */
int smx_plan_project(smx_plan *plan, int nmodes, const double *modes, double *S_mon, double *S_dft);

/*
 * Extension: adjoint gradient of a port objective (FP32 plans without lossy or metal
 * cells).  No counterpart in the reference.
 *   J = sum_{v = x, y} sum_q |D_v(q)|^2,  D_v(q) = sum_k e^{i*w*(k+1.5)} E_v(q, zT) after step k
 * over steps k = 0 .. nsteps-1 of the plan's source loop, from zero fields.
 * smx_adjoint_forward runs the forward pass, storing the full state every ckpt_every
 * steps, and returns J.  smx_adjoint_backward runs the discrete adjoint in reverse
 * time -- two transposed steps per launch, bit for bit what one step per launch gives
 * (SMX_ADJ_TWO=0) -- recomputing each checkpoint segment forward (two steps per pass,
 * writing E of the design box after each) and writes grad[3][nbox] = dJ/d eps_c
 * (c = x, y, z) for the cells of the design box box = {x0, x1, y0, y1, z0, z1}
 * (half-open, x fastest), which must hold only 'n' cells.  Both leave the plan's fields
 * at the last recomputed state; the recomputation leaves its monitors alone.
 * ckpt_every = 0 trades memory for the recomputation: the forward pass stores E of the
 * design box after every step (12*nbox*(nsteps+1) bytes of device memory instead of 48*n
 * per checkpoint) and the backward pass recomputes nothing.
 */
int smx_adjoint_forward(smx_plan *plan, long long nsteps, int ckpt_every, int zT, double w, const int box[6],
                        double *J);
int smx_adjoint_backward(smx_plan *plan, double *grad);

/*
 * Autotuner of the FP32 step: times the two in-place sweeps and the fused step in each
 * variant -- tiles of 32x8, 64x4 and 16x16 cells of (j0, j1) staged through registers,
 * and tiles of 32x8 and 16x16 staged by asynchronous copies to LDS -- with each z chunk
 * (8, 16, 32, 64 planes) for trial_steps steps on this plan's grid, keeps the fastest,
 * and restores fields and monitors.  choice (may be NULL) receives {tile x, tile y,
 * z chunk, variant, z balance}: variant 0 = in-place sweeps (sizes 0, also for FP64
 * plans), 1 = register-staged fused step, 2 = LDS-staged fused step, 3 = two steps per
 * pass (grids without lossy/metal cells; used for pairs of steps of the source loop,
 * single steps take the best one-step kernel; with a DFT set, only the variants that
 * sample the DFT planes of both steps themselves are considered); z balance (variant 3
 * only) 0 = z chunks of equal size, k > 0 = chunks balanced with a z-PML stage counted
 * 1 + k/2 times.  trial_steps = 0 only reports.  No counterpart in the reference.
 */
int smx_plan_autotune(smx_plan *plan, int trial_steps, int choice[5]);

/* Names of the kernels one step (one pass, with two steps per pass) of the plan launches
 * with its current kernel choice and DFT setting, " + "-separated, as profilers list
 * them (lossy/metal passes and the DFT fold not included); a static string.  No
 * counterpart in the reference. */
const char *smx_plan_kernels(const smx_plan *plan);

/* The plan's HIP stream (a hipStream_t), for callers that order or time their own
 * work against the plan's kernels.  Valid until smx_plan_destroy. */
void *smx_plan_stream(const smx_plan *plan);

/* ------------------------------------------------------------------------------------
 * Host setup helpers: the reference's setup functions, FP64, same results.
 * ---------------------------------------------------------------------------------- */

/* set_tau (F:515-524): returns N_2pi; writes 1/tau_R, 1/tau_I and Dt. */
int smx_set_tau(double omega, double min_Dt, double *invtau_R, double *invtau_I, double *Dt);
/* set_env (F:452-475); NAN where the reference exits. */
double smx_set_env(int N_2pi, const int time_period[3], int j_half, int j_time, int j_period);
/* ExpN table (F:93-99): complex[2*N_2pi]. */
void smx_expn_table(int N_2pi, double *ExpN);
/* set_sigma (F:858-926) into a zeroed sigma[2*(L0+L1+L2)]. */
void smx_set_sigma(const int L[4], const double *mu, const double *eps, double *sigma);
/* remove_sigma (F:822-853): 0, or 1/2 where the reference exits on mu/eps. */
int smx_remove_sigma(const int L[4], double *mu, double *eps, const double *sigma);
/* set_address (F:188-233): returns Lnotn. */
int smx_set_address(const int L[4], const double *sigma, const double *mu, const double *eps,
                    char *Cont, int *address_notn);
/* Coefficient build (F:106-143) into zeroed arrays of 3*(n+Lnotn) doubles. */
void smx_build_coeffs(const int L[4], const double *mu, const double *eps, const double *sigma,
                      const char *Cont, const int *address_notn, int Lnotn, double invtau_R,
                      double invtau_I, double *Ah, double *Bh, double *Ae, double *Be);

#ifdef __cplusplus
}
#endif

#endif /* SMX_FDTD_H */
