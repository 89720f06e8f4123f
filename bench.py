#!/usr/bin/env python3
"""Headline benchmark: split-field FDTD cell-updates/s on one MI355X.

Workload (one "step" = one full leapfrog step, H and E, over the whole grid): a
curvilinear 160 x 160 x 320 Yee grid (scale factors h_j = 1 + 0.3 g_j from a seeded
sum of 8 low-order Fourier modes), a Si-like core eps = 12.1 in eps = 2.1 cladding,
PEC on ~5% of the scatterer's E edges as seeded slabs, 10-cell PML on the z faces
and PEC walls on x/y (the layout the reference's monitors accept), a modal sheet
source with the reference's sin^2/CW envelope, and the reference's running DFT
monitors on two port planes.  Inputs are synthetic and resident in HBM before the
timed region.

Prints one JSON line: FP32 Gcell-updates/s (value; the step kernel chosen by the
plan's autotuner), the roofline of that kernel against MI355X HBM peak, the plain FP64
path on the same grid, the FP32-vs-FP64 |S| error (a separate run of 10,000 steps from
zero fields, the envelope's ramp then CW, whatever --steps is), and the reference's own main_calc
timed on this host's cores (oracle/_ref; the restatement in oracle/ where that is not
built).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "fdtd_f32_step_160x160x320_pmc.txt")
METRIC = "Gcell-updates/s + % HBM peak, FP32 vs FP64 GPU path at <=1e-3 |S| error"


def bytes_per_cell(phi, field_bytes, coef_bytes):
    """Algorithmic bytes of one cell-update: every field component read and written
    once (12 words), split/aux slots in a fraction phi of cells (12*phi words), one
    coefficient word per component (6 words)."""
    return field_bytes * (12 + 12 * phi) + 6 * coef_bytes


def measured_traffic(kernels, steps_per_pass=1):
    """HBM bytes of one step from the committed rocprofv3 FETCH_SIZE and WRITE_SIZE
    averages (KiB per dispatch) of the 160x160x320 grid, each scaled by what the counter
    reads for a plain device copy of known size.  `kernels` lists the launches of one
    pass (a kernel launched twice per pass appears twice; its average is over all its
    dispatches), and a pass advances `steps_per_pass` steps.  None when a kernel is not
    listed there."""
    try:
        with open(PMC_SUMMARY) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    scale, table = None, {}
    for ln in lines:
        w = ln.split()
        if not w or w[0].startswith("#"):
            continue
        if w[0] == "scale":
            scale = (float(w[2]), float(w[4]))
        else:
            table[w[0]] = (float(w[1]), float(w[2]))
    if scale is None or any(k not in table for k in kernels):
        return None  # the committed profile does not list the kernel this run chose: stale
    return sum((table[k][0] * scale[0] + table[k][1] * scale[1]) * 1024 for k in kernels) / steps_per_pass


def copy_bandwidth(torch, mib=1024, reps=5):
    """Achievable HBM bandwidth: a plain device-to-device copy of `mib` MiB, timed with
    HIP events; bytes read + written per second, GB/s."""
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    gbs = 2 * 4 * n * reps / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def overlap(prep, H, E):
    """Projection of monitor phasors (H, E: complex[2*L0*L1], (x, y) interleaved) on the
    source profile, normalised by the profile norm: the modal overlap S_calc forms
    (F:437-445) with the incident profile as the mode vector."""
    import numpy as np
    c = prep.case
    w = np.concatenate([c.Hinc, c.Einc])
    v = np.concatenate([H, E])
    return complex(np.vdot(w, v) / np.vdot(w, w).real)


ACC_STEPS = 10000  # steps of the accuracy run (the 160x160x320 configuration's run length), ramp included


def run_from_zero(plan, dft_planes, dft_w, chunk, steps):
    """Zero the plan's fields, monitors and DFT, run steps [0, steps) of the source loop
    and return (monitors, DFT or None, envelope sums)."""
    import numpy as np
    if dft_planes:
        plan.set_dft(dft_planes, dft_w, chunk=chunk)
    z = np.zeros(2 * plan.grid.plane, dtype=np.complex128)
    plan.write_monitors(z, z, z, z)
    plan.write_fields(np.zeros(plan.flen), np.zeros(plan.flen))
    sums = plan.run_steps(0, steps)
    return plan.read_monitors(), (plan.read_dft() if dft_planes else None), sums, plan.read_fields()


def s_comparison(prep, r32, r64, dft_planes):
    """FP32 against FP64 on the |S| values FDTD would print: the monitor phasors and every
    DFT (plane, frequency), projected on the source profile and divided by the envelope
    sum of the steps run (sum00, F:159-168), which makes a partial run's |S| the size of a
    whole run's.  Raises when an FP64 value is zero (a monitor the wave has not reached
    compares nothing)."""
    import numpy as np
    (m32, d32, _, f32), (m64, d64, sums, f64) = r32, r64
    env = float(sums[0])
    vals = {}
    for name, k in (("bottom", 0), ("top", 1)):
        if prep.case.iB >= 0:
            vals[name] = (overlap(prep, m32[2 * k], m32[2 * k + 1]), overlap(prep, m64[2 * k], m64[2 * k + 1]))
    if dft_planes:
        def phasors(d, m, f):
            H = np.empty(2 * prep.case.plane, dtype=np.complex128)
            E = np.empty_like(H)
            H[0::2], H[1::2], E[0::2], E[1::2] = d[m, 0, f], d[m, 1, f], d[m, 2, f], d[m, 3, f]
            return H, E
        for m in range(len(dft_planes)):
            for f in range(d64.shape[2]):
                vals[f"dft_plane{m}_f{f}"] = (overlap(prep, *phasors(d32, m, f)), overlap(prep, *phasors(d64, m, f)))
    errs = {}
    for key, (a, b) in vals.items():
        a, b = abs(a) / env, abs(b) / env
        if not b > 0:
            raise RuntimeError(f"|S| comparison: FP64 value '{key}' is zero; the wave has not reached it")
        errs[key] = (abs(a - b), b)
    worst = max(errs, key=lambda k: errs[k][0] / errs[k][1])
    return dict(S_rel_err_max=max(e / b for e, b in errs.values()), S_rel_err_at=worst,
                S_abs_err_max=max(e for e, _ in errs.values()), S_abs_min=min(b for _, b in errs.values()),
                S_abs_max=max(b for _, b in errs.values()), compared=len(errs),
                monitors={k: dict(abs_fp64=errs[k][1], abs_err=errs[k][0]) for k in ("bottom", "top") if k in errs},
                envelope_sum=env,
                fields_rel_l2=float(np.linalg.norm(np.concatenate([f32[0] - f64[0], f32[1] - f64[1]])) /
                                    np.linalg.norm(np.concatenate(f64))))


def build_workload(small=False, box=False):
    from smatranem_amd import casegen
    L, pml = ((32, 32, 48), 4) if small else ((160, 160, 320), 10)
    return casegen.curvilinear_waveguide_case(L, pml=pml, pml_xy=pml if box else 0)


def time_plan(plan, steps, warmup, torch):
    stream = torch.cuda.ExternalStream(plan.stream)
    plan.run_steps(0, warmup)
    plan.sync()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    t0 = time.perf_counter()
    plan.run_steps(warmup, steps)
    plan.flush()  # the DFT and monitor samples still in the ring are folded inside the region
    ev1.record(stream)
    plan.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3, ev0.elapsed_time(ev1)


def box_run(args, torch, K, W):
    """The headline grid with 10-cell PML on all six faces (the
    reference monitors off: every plane holds PML cells) and the 16-frequency DFT on the
    two port planes, autotuned, max(K/5, 100) steps timed."""
    import numpy as np
    from smatranem_amd import casegen, fdtd
    bcase = build_workload(args.small, box=True)
    bprep = casegen.prepare(bcase)
    bpml = 4 if args.small else 10
    bplanes = [] if args.no_dft else [bpml + 2, int(bcase.L[2]) - 1 - bpml - 2]
    dft_w = [float(x) for x in 2 * np.pi / bprep.N * np.linspace(0.85, 1.15, 16)]
    with bprep.plan(fdtd.F32) as pb:
        if bplanes:
            pb.set_dft(bplanes, dft_w, chunk=args.dft_chunk)
        pb.autotune(20)
        bkern = pb.kernels
        kb = max(K // 5, 100)
        wb, evb = time_plan(pb, kb, min(W, 100), torch)
    bphi = bprep.pml_fraction
    bb = bytes_per_cell(bphi, 4, 4)
    box = dict(workload="same grid, 10-cell PML on all six faces, reference monitors off, "
                        + ("16-frequency DFT on planes %s" % bplanes if bplanes else "no DFT"),
               kernel=bkern, steps=kb, value=round(bcase.n / (wb / kb * 1e-3) / 1e9, 4),
               pml_fraction=round(bphi, 4), bytes_per_cell=round(bb, 2),
               kernel_ms_per_step=round(evb / kb, 5),
               roofline_frac=round(bb * bcase.n / (evb / kb * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
    print(f"[bench] box-PML fp32: {box['value']:.2f} Gcell/s, {box['roofline_frac']:.3f} of peak ({bkern})",
          file=sys.stderr, flush=True)
    return box


def metal_run(args, torch, K, W):
    import numpy as np
    """The headline grid with an eps-metal block (eps = -20 + 2i, class 'e', 16 x 16 x 16
    cells in the cladding beside the core, 1/2000 of the grid): the reference monitors and
    the 16-frequency DFT on the two port planes as in the headline, autotuned, max(K/5,
    100) steps timed; also the best one-step kernel on the same grid."""
    import os
    from smatranem_amd import casegen, fdtd
    L, pml = ((32, 32, 48), 4) if args.small else ((160, 160, 320), 10)
    box = (12, 16, 12, 16, 20, 24) if args.small else (100, 116, 72, 88, 150, 166)
    mcase = casegen.curvilinear_waveguide_case(L, pml=pml, metal_box=(box, -20.0 + 2.0j))
    mprep = casegen.prepare(mcase)
    planes = [] if args.no_dft else [mcase.iB, mcase.iT]
    dft_w = [float(x) for x in 2 * np.pi / mprep.N * np.linspace(0.85, 1.15, 16)]
    kb = max(K // 5, 100)
    out = {}
    for label, two in (("two_step", None), ("one_step", "off")):
        if two:
            os.environ["SMX_F32_TWO"] = two  # the autotuner then times one-step kernels only
        try:
            with mprep.plan(fdtd.F32) as pm:
                if planes:
                    pm.set_dft(planes, dft_w, chunk=args.dft_chunk)
                choice = pm.autotune(20)
                assert (choice[3] == 3) == (two is None), (label, choice)
                kern = pm.kernels
                wb, evb = time_plan(pm, kb, min(W, 100), torch)
                H, E = pm.read_fields()
                finite = bool(np.isfinite(H).all() and np.isfinite(E).all())
        finally:
            os.environ.pop("SMX_F32_TWO", None)
        out[label] = dict(kernel=kern, value=round(mcase.n / (wb / kb * 1e-3) / 1e9, 4),
                          kernel_ms_per_step=round(evb / kb, 5), fields_finite=finite)
    bphi = mprep.pml_fraction
    bb = bytes_per_cell(bphi, 4, 4)
    two = out["two_step"]
    res = dict(workload="same grid, eps-metal block %s (eps -20+2i, class 'e'), z PML, reference monitors, %s"
                        % (list(box), "16-frequency DFT on planes %s" % planes if planes else "no DFT"),
               special_cells=int(np.count_nonzero(mprep.Cont == ord("e"))), steps=kb, kernel=two["kernel"],
               value=two["value"], kernel_ms_per_step=two["kernel_ms_per_step"],
               roofline_frac=round(bb * mcase.n / (two["kernel_ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
               fields_finite=two["fields_finite"], one_step=out["one_step"])
    print(f"[bench] metal fp32: {res['value']:.2f} Gcell/s ({res['kernel']}), one step per pass "
          f"{out['one_step']['value']:.2f} ({out['one_step']['kernel']})", file=sys.stderr, flush=True)
    return res


def point_run(torch, steps=2000):
    """The broadband point-source workload: a 48 x 48 x 64 unit grid, vacuum with a centred
    16 x 16 x 24 block of eps 12.1, 8-cell cubic PML on every face, a Gaussian-modulated
    E_z point current source (centre 0.1 cycles per cell, 1/e spectral half-width 0.05) at
    (24, 24, 14), `steps` steps at 0.99 of the 3-D CFL limit, a DFT at 4 frequencies on
    plane 50.  FP64 (the plain path) and FP32 (autotuned, one step per pass), timed over
    the whole run; the FP32 DFT against the FP64 one.  A cache-resident grid: the times
    are launch-bound, not a roofline measurement."""
    import dataclasses
    import numpy as np
    from smatranem_amd import casegen, fdtd
    case = casegen.point_source_case()
    prep = casegen.prepare(case)
    iR, iI, dt, amp = casegen.point_source_drive(case, steps)
    prep = dataclasses.replace(prep, invtau_R=iR, invtau_I=iI, dt=dt)
    freqs = (0.085, 0.095, 0.105, 0.115)
    ws = [2 * np.pi * f * dt for f in freqs]
    out = {}
    for prec in (fdtd.F64, fdtd.F32):
        with prep.plan(prec) as p:
            p.set_dft([50], ws, chunk=64)
            p.set_point_source(2, (24, 24, 14), amp)
            if prec == fdtd.F32:
                p.autotune(10)
            kern = p.kernels
            wall, ev = time_plan(p, steps, 0, torch)
            out[prec] = (wall, ev, kern, p.read_dft().reshape(4, len(ws), -1))
    d64, d32 = out[fdtd.F64][3], out[fdtd.F32][3]
    errs = [float(np.abs(d32[:, f] - d64[:, f]).max() / np.abs(d64[:, f]).max()) for f in range(len(ws))]
    g = lambda prec: round(case.n * steps / (out[prec][0] * 1e-3) / 1e9, 4)  # noqa: E731
    res = dict(workload="48x48x64 unit grid, eps 12.1 block in vacuum, 8-cell PML on every face, Gaussian-modulated "
                        "E_z point source (f0 0.1, bw 0.05), %d steps at 0.99 CFL (Dt %.6f), 4-frequency DFT on "
                        "plane 50" % (steps, dt),
               cells=case.n, steps=steps, fp64_gcells=g(fdtd.F64), fp64_kernel=out[fdtd.F64][2],
               fp32_gcells=g(fdtd.F32), fp32_kernel=out[fdtd.F32][2],
               dft_freqs=list(freqs), dft_rel_err_fp32_vs_fp64=errs)
    print(f"[bench] point source: FP64 {res['fp64_gcells']} Gcell/s, FP32 {res['fp32_gcells']} Gcell/s, "
          f"FP32 DFT error {max(errs):.2e} of its largest value", file=sys.stderr, flush=True)
    return res


SMALL_BATCH = dict(dims=((24, 48), (24, 48), (32, 64)), pml=4, seed=2026)


def batch_run(count, steps, torch, small=False, dims=((64, 192), (64, 192), (64, 256)), pml=8, seed=2025,
              ports=False):
    """Independent runs of seeded sizes (L0, L1, L2 uniform in `dims`, dielectric boxes,
    `pml`-cell PML on every face), FP32, each plan autotuned: with `ports`, count/12
    structures each excited by 6 Hermite-Gaussian port modes in 2 polarisations
    (casegen.port_batch_cases), else `count` structures of one excitation each.  After the
    envelope's ramp (untimed), `steps` steps of all plans are timed four ways: the fused
    batch mode (one launch of two_step_multi per two-step pass of all plans; `value`),
    smx_plan_run_batch (one captured HIP graph of every plan's segment, replayed), the
    same batch with every step launched from the host (streams mode), and the plans one
    after another."""
    import dataclasses
    from smatranem_amd import casegen, fdtd
    if ports:
        runs = casegen.port_batch_cases(max(count // 12, 1), 6, 2, seed=seed, pml=pml, dims=dims)
        base = {}
        preps = []
        for si, m, pol, c in runs:
            if si not in base:
                base[si] = casegen.prepare(c)
            preps.append(dataclasses.replace(base[si], case=c))
        cases = [pr.case for pr in preps]
    else:
        cases = casegen.batch_cases(count, seed=seed, pml=pml, dims=dims)
        if small:
            cases = [casegen.waveguide_case(24 + 8 * i, 20, 40, outer=8, pml_xy=4, pml_z=4, seed=i, blocks=2,
                                            time_period=(5, 100000, 100010)) for i in range(count)]
        preps = [casegen.prepare(c) for c in cases]
    plans = [pr.plan(fdtd.F32) for pr in preps]
    warm = max(pr.N * (pr.case.time_period[0] + 1) for pr in preps)

    def sync_all():
        for p in plans:
            p.sync()
        torch.cuda.synchronize()

    def timed(fn):
        sync_all()
        t0 = time.perf_counter()
        fn()
        sync_all()
        return time.perf_counter() - t0
    info, info_f = {}, {}
    try:
        for p in plans:
            p.autotune(10)
        fdtd.run_batch(plans, 0, warm, mode=fdtd.BATCH_STREAMS)
        k = warm
        fdtd.run_batch(plans, k, 20, mode=fdtd.BATCH_FUSED)  # first launch of the batch kernel, untimed
        k += 20
        t_fused = timed(lambda: fdtd.run_batch(plans, k, steps, mode=fdtd.BATCH_FUSED, info=info_f))
        assert info_f["graph_launches"] == steps // 2, "the batch did not take the fused launch"
        k += steps
        t_graph = timed(lambda: fdtd.run_batch(plans, k, steps, info=info))
        k += steps
        t_streams = timed(lambda: fdtd.run_batch(plans, k, steps, mode=fdtd.BATCH_STREAMS))
        k += steps

        def seq():
            for p in plans:
                p.run_steps(k, steps)
        t_seq = timed(seq)
    finally:
        for p in plans:
            p.close()
    cells = sum(c.n for c in cases)
    alg = sum(c.n * bytes_per_cell(pr.pml_fraction, 4, 4) for c, pr in zip(cases, preps))
    what = ("%d structures x 6 Hermite-Gaussian port modes x 2 polarisations" % max(count // 12, 1) if ports
            else "%d independent runs" % count)
    return dict(workload="%s, L0 in [%d, %d], L1 in [%d, %d], L2 in [%d, %d], dielectric boxes, "
                         "%d-cell PML on all faces, FP32, smx_plan_run_batch_mode SMX_BATCH_FUSED (one "
                         "two_step_multi launch per two-step pass of all plans)"
                         % (what, *dims[0], *dims[1], *dims[2], pml),
                grids=[list(c.L[:3]) for c in cases], cells=cells, steps=steps, first_step=warm,
                value=round(cells * steps / t_fused / 1e9, 4), unit="Gcell-updates/s",
                roofline_frac=round(alg * steps / t_fused / 1e9 / HBM_PEAK_GBS, 4),
                fused_launches=info_f.get("graph_launches"),
                graph_replay=round(cells * steps / t_graph / 1e9, 4),
                graph_launches=info.get("graph_launches"),
                streams_mode=round(cells * steps / t_streams / 1e9, 4),
                one_after_another=round(cells * steps / t_seq / 1e9, 4))


def adjoint_run(torch, small=False, fd_cells=0, fd_seed=None, steps=None):
    """The gradient workload: a 96 x 96 x 128 design region of per-cell eps in [2.1, 12.1]
    inside a 192 x 192 x 256 grid with 10-cell PML on every face; forward 4000 FP32 steps
    with the full state checkpointed every 200 steps (smx_adjoint_forward), then the
    reverse-time adjoint with recomputation (smx_adjoint_backward).  Objective: DFT power
    of the tangential E at the source frequency on a plane 3 cells below the top PML.
    fd_cells > 0 also checks dJ/d eps at that many cells against central differences
    (delta eps = 1e-3) of the FP64 forward: the largest-gradient cells, or with fd_seed
    cells (and components) drawn uniformly from the design box with that seed.  Each
    check reports the error relative to the cell's FD value and to the largest |dJ/d eps|
    (FP32 rounding sets a floor on the second)."""
    import numpy as np
    from smatranem_amd import casegen, fdtd
    if small:
        L, pml, K, C = (48, 48, 64), 6, 400, 50
    else:
        L, pml, K, C = (192, 192, 256), 10, 4000, 200
    if steps:
        K, C = steps, min(C, steps)
    case, box = casegen.adjoint_case(L, pml=pml)
    prep = casegen.prepare(case)
    zT = L[2] - pml - 3
    w = 2 * np.pi / prep.N
    n = case.n
    nbox = (box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4])
    modes = {}
    g = None
    with prep.plan(fdtd.F32) as p:
        p.autotune(10)
        kern = p.kernels
        # checkpoints every C steps with recomputation; then the design box's E stored
        # after every step (ckpt_every = 0), no recomputation
        for ck in (C, 0):
            p.sync()
            t0 = time.perf_counter()
            J = p.adjoint_forward(K, ck, zT, w, box)
            t1 = time.perf_counter()
            gk = p.adjoint_backward()
            t2 = time.perf_counter()
            if g is None:
                g = gk
            mem = 12 * 4 * n * ((K + ck - 1) // ck) if ck else 3 * 4 * nbox * (K + 1)
            modes["checkpoint_recompute" if ck else "box_history"] = dict(
                ckpt_every=ck, forward_s=round(t1 - t0, 3), backward_s=round(t2 - t1, 3),
                forward_gcells=round(n * K / (t1 - t0) / 1e9, 3), backward_gcells=round(n * K / (t2 - t1) / 1e9, 3),
                stored_GB=round(mem / 1e9, 2), J=J,
                grad_rel_diff_vs_checkpoint=float(np.linalg.norm(gk - g) / np.linalg.norm(g)))
            print(f"[bench] adjoint {'every %d' % ck if ck else 'box history'}: forward "
                  f"{modes[list(modes)[-1]]['forward_gcells']} Gcell/s, backward "
                  f"{modes[list(modes)[-1]]['backward_gcells']} Gcell/s", file=sys.stderr, flush=True)
    ref = modes["checkpoint_recompute"]
    out = dict(workload="adjoint gradient, %dx%dx%d grid, %d-cell PML, design box %s (eps U[2.1, 12.1]), "
                        "%d FP32 steps, checkpoints every %d" % (L[0], L[1], L[2], pml, list(box), K, C),
               cells=n, design_cells=nbox, steps=K, ckpt_every=C, step_kernel=kern,
               forward_s=ref["forward_s"], backward_s=ref["backward_s"], forward_gcells=ref["forward_gcells"],
               backward_gcells=ref["backward_gcells"], checkpoint_GB=ref["stored_GB"], J=ref["J"],
               grad_l2=float(np.linalg.norm(g)), modes=modes)
    if fd_cells > 0:
        def J64(eps):
            c2 = casegen.Case(**{**case.__dict__, "eps": eps})
            pr = casegen.prepare(c2)
            with pr.plan(fdtd.F64) as q:
                q.set_dft([zT], [w], chunk=32)
                q.run_steps(0, K)
                D = q.read_dft()[0]
            return float(np.sum(np.abs(D[2:4, 0, :]) ** 2))
        x0, x1, y0, y1, z0, z1 = box
        if fd_seed is None:
            order = np.argsort(np.abs(g).reshape(-1))[::-1][:fd_cells]
        else:
            order = np.random.default_rng(fd_seed).choice(g.size, size=fd_cells, replace=False)
        gmax = float(np.abs(g).max())
        checks = []
        for idx in order:
            c, zz, yy, xx = (int(v) for v in np.unravel_index(idx, g.shape))
            cell = (x0 + xx) + L[0] * ((y0 + yy) + L[1] * (z0 + zz))
            h = 1e-3
            ep, em = case.eps.copy(), case.eps.copy()
            ep[3 * cell + c] += h
            em[3 * cell + c] -= h
            fd = (J64(ep) - J64(em)) / (2 * h)
            err = abs(float(g[c, zz, yy, xx]) - fd)
            checks.append(dict(cell=[x0 + xx, y0 + yy, z0 + zz], comp=c, adjoint=float(g[c, zz, yy, xx]), fd=fd,
                               rel_err=err / abs(fd), err_rel_to_max=err / gmax))
            print(f"[bench] adjoint vs FD at {checks[-1]['cell']} c={c}: {checks[-1]['rel_err']:.2e} of the cell's "
                  f"value, {checks[-1]['err_rel_to_max']:.2e} of max |dJ/d eps|", file=sys.stderr, flush=True)
        out["fd_check"] = checks
        out["fd_cells"] = "largest |dJ/d eps|" if fd_seed is None else "seeded uniform draw, seed %d" % fd_seed
        out["fd_max_rel_err"] = max(ch["rel_err"] for ch in checks)
        out["fd_max_err_rel_to_max"] = max(ch["err_rel_to_max"] for ch in checks)
    return out


def large_run(torch, steps, depth=256, small=False):
    """The large-grid workload: 768 x 768 x `depth` cells (1024: the full 604 M-cell grid;
    256: a quarter of it), curvilinear scale factors, 12 dielectric layers with a PEC
    grating, 12-cell PML on every face, an 8-frequency DFT (FP64 accumulators) on two
    planes.  FP32 (autotuned) over `steps` steps from zero fields; up to 200 M cells also
    FP64 over the same steps, with the |S| error of FP32 against FP64 from the DFT phasors
    projected on the source profile (beyond, the FP64 plan's host coefficient arrays alone
    would take 70 GB and more)."""
    import numpy as np
    from smatranem_amd import casegen, fdtd
    L, pml = ((96, 96, 64), 6) if small else ((768, 768, depth), 12)
    t0 = time.perf_counter()

    def say(msg):
        print(f"[bench] large grid {L}: {msg} ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    case = casegen.layered_stack_case(L, pml=pml, progress=say)
    prep = casegen.prepare(case, progress=say)
    say("setup done")
    planes = [pml + 3, L[2] - pml - 4]
    ws = [float(x) for x in 2 * np.pi / prep.N * np.linspace(0.9, 1.1, 8)]
    out = {}
    with_fp64 = case.n <= 200_000_000
    for prec in (fdtd.F32, fdtd.F64) if with_fp64 else (fdtd.F32,):
        with prep.plan(prec) as p:
            say("%s plan created" % ("FP32" if prec == fdtd.F32 else "FP64"))
            p.set_dft(planes, ws, chunk=64)
            kern = p.kernels
            if prec == fdtd.F32:
                p.autotune(6)
                kern = p.kernels
            wall, ev = time_plan(p, steps, 0, torch)
            out[prec] = (ev / steps, p.read_dft(), kern)
        print(f"[bench] large grid {'FP32' if prec == fdtd.F32 else 'FP64'}: {out[prec][0]:.3f} ms/step",
              file=sys.stderr, flush=True)
    n, phi = case.n, prep.pml_fraction
    g32 = n / (out[fdtd.F32][0] * 1e-3) / 1e9
    res = dict(workload="layered stack %dx%dx%d: curvilinear scale factors, 12 dielectric layers, PEC "
                        "grating, %d-cell PML on all faces, 8-frequency DFT on 2 planes" % (*L, pml),
               cells=n, pml_fraction=round(phi, 4), steps=steps, kernel=out[fdtd.F32][2],
               fp32_gcells=round(g32, 3), fp32_ms_per_step=round(out[fdtd.F32][0], 4),
               fp32_roofline_frac=round(bytes_per_cell(phi, 4, 4) * g32 / HBM_PEAK_GBS, 4),
               fp32_device_GB=round(30 * 4 * n / 1e9, 1))
    if not with_fp64:
        return res
    errs = []
    for m in range(len(planes)):
        for f in range(len(ws)):
            def ph(d):
                H = np.empty(2 * case.plane, dtype=np.complex128)
                E = np.empty_like(H)
                H[0::2], H[1::2], E[0::2], E[1::2] = d[m, 0, f], d[m, 1, f], d[m, 2, f], d[m, 3, f]
                return H, E
            s32, s64 = overlap(prep, *ph(out[fdtd.F32][1])), overlap(prep, *ph(out[fdtd.F64][1]))
            errs.append((abs(abs(s32) - abs(s64)), abs(s64)))
    g64 = n / (out[fdtd.F64][0] * 1e-3) / 1e9
    return dict(res, fp64_gcells=round(g64, 3),
                fp64_roofline_frac=round(bytes_per_cell(phi, 8, 8) * g64 / HBM_PEAK_GBS, 4),
                speedup_fp32_over_fp64=round(g32 / g64, 3),
                S_abs_err_max=max(e for e, _ in errs), S_abs_fp64_max=max(r for _, r in errs),
                S_err_rel_to_max=max(e for e, _ in errs) / max(max(r for _, r in errs), 1e-300))


def cpu_model():
    """Model name, family/model/stepping and logical CPU counts of this host."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                k, _, v = ln.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                    info[k] = v.strip()
                if k == "processor":
                    info["logical_cpus"] = info.get("logical_cpus", 0) + 1
    except OSError:
        return None
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return ", ".join(f"{k} {v}" for k, v in info.items())


def cpu_baseline(prep, seconds, all_cores=False):
    """The reference's main_calc (compiled from its source, oracle/_ref) or, where that
    is not built, the oracle's restatement, single-threaded on this host, over whole
    steps of the same grid until ~`seconds` have elapsed.  all_cores: the oracle's
    restatement with its sweeps on OpenMP threads (OMP_NUM_THREADS of them, else one per
    host core), bit for bit the same fields; None when that build is absent."""
    import numpy as np
    import oracle
    c = prep.case
    N, iR, iI, _ = oracle.set_tau(c.omega, c.min_dt / c.dt_factor)
    st = oracle.setup_from_media(c.L, c.mu, c.eps, iR=iR, iI=iI)
    threads = 1
    if all_cores:
        omp = oracle.load_oracle_omp()
        if omp is None:
            return None
        kind, fn = "port", omp.orc_main_calc
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    else:
        ref = oracle.load_ref()
        kind = "reference" if ref is not None else "port"
        fn = ref.main_calc if ref is not None else oracle.load_oracle().orc_main_calc
    n, plane = c.n, c.plane
    flen = 3 * (n + st["Lnotn"] + 1)
    H, E = np.zeros(flen), np.zeros(flen)
    mons = [np.zeros(4 * plane) for _ in range(4)]
    ExpN = oracle.expn_table(N)
    tp = c.time_period
    Hinc = np.ascontiguousarray(c.Hinc).view(np.float64)
    Einc = np.ascontiguousarray(c.Einc).view(np.float64)
    steps = 0
    t0 = time.perf_counter()
    while True:
        t = steps % N
        P = steps // N
        e = ExpN[4 * t:4 * t + 4]
        tn = (t + 1) % N
        fE = 2.0 * oracle.set_env(N, tp, 1, t, P)
        fH = 2.0 * oracle.set_env(N, tp, 0, t + 1, P)
        Ei = fE * (e[2] * Einc[0::2] + e[3] * Einc[1::2])
        Hi = fH * (ExpN[4 * tn] * Hinc[0::2] + ExpN[4 * tn + 1] * Hinc[1::2])
        oracle.main_calc(fn, c.L, st["Ah"], st["Bh"], st["Ae"], st["Be"], c.iBc, H, E,
                         (c.iB, c.iT, c.iI, c.inc_mode), np.ascontiguousarray(Hi),
                         np.ascontiguousarray(Ei), ExpN[4 * tn:4 * tn + 4].copy(), *mons,
                         st["Cont"], st["addr"])
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return dict(value=round(n * steps / el / 1e9, 6), unit="Gcell-updates/s", cores=threads, kind=kind,
                sample=f"{steps} full steps of the same {c.L[0]}x{c.L[1]}x{c.L[2]} grid, "
                       + (f"{threads} OpenMP threads" if all_cores else "single-threaded") + f", {el:.1f} s",
                cpu_model=cpu_model())


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg")
    ap.add_argument("--no-box", action="store_true", help="skip the all-faces-PML variant")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch of independent runs")
    ap.add_argument("--no-metal", action="store_true", help="skip the grid with an eps-metal block")
    ap.add_argument("--batch-count", type=int, default=24)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint gradient workload")
    ap.add_argument("--no-dft", action="store_true", help="leave the 16-frequency port DFT off")
    ap.add_argument("--dft-chunk", type=int, default=256, help="steps between folds of the DFT samples (1..256)")
    ap.add_argument("--large", type=int, default=0, metavar="STEPS",
                    help="also run the 768x768xDEPTH layered-stack grid for STEPS steps (FP32, and FP64 up to "
                         "200 M cells)")
    ap.add_argument("--large-depth", type=int, default=256, metavar="DEPTH",
                    help="z cells of the large grid (1024: 768x768x1024, 604 M cells)")
    ap.add_argument("--adjoint-steps", type=int, default=0, metavar="K",
                    help="steps of the adjoint workload (default 4000, checkpoints every 200)")
    ap.add_argument("--adjoint-fd", type=int, default=0, metavar="CELLS",
                    help="check the adjoint gradient at CELLS cells against FP64 central differences")
    ap.add_argument("--adjoint-fd-seed", type=int, default=None, metavar="SEED",
                    help="draw the --adjoint-fd cells uniformly from the design box with this seed")
    ap.add_argument("--small", action="store_true", help="32x32x48 grid (quick check)")
    ap.add_argument("--no-point", action="store_true", help="skip the broadband point-source grid")
    ap.add_argument("--only", choices=["box", "metal", "batch", "batch-small", "adjoint", "point"], default=None,
                    help="run only this leg (for profiling) and print its JSON object")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (scale for rocprofv3 byte counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: runs on one GPU only; the multi-GPU path (z slabs with a halo exchange "
              "per half step) is not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from smatranem_amd import casegen, fdtd

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n,
                          "copies": 5}), flush=True)
        return
    K, W = args.steps, args.warmup
    if args.only:
        if args.only == "box":
            out = box_run(args, torch, K, W)
        elif args.only == "metal":
            out = metal_run(args, torch, K, W)
        elif args.only == "batch":
            out = batch_run(args.batch_count, args.batch_steps, torch, args.small, ports=not args.small)
        elif args.only == "batch-small":
            out = batch_run(args.batch_count, 4 * args.batch_steps, torch, args.small, **SMALL_BATCH)
        elif args.only == "point":
            out = point_run(torch)
        else:
            out = adjoint_run(torch, args.small, args.adjoint_fd, args.adjoint_fd_seed, args.adjoint_steps)
        print(json.dumps({"only": args.only, "result": out}), flush=True)
        return

    t_setup = time.perf_counter()
    case = build_workload(args.small)
    prep = casegen.prepare(case)
    total = W + K
    assert total <= (case.time_period[2]) * prep.N
    n = case.n
    phi = prep.pml_fraction
    # 16 frequencies around the source's on the two port planes (FP64 accumulators)
    dft_planes = [] if args.no_dft else [case.iB, case.iT]
    dft_w = [float(x) for x in 2 * np.pi / prep.N * np.linspace(0.85, 1.15, 16)]
    dft_chunk = args.dft_chunk
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] grid {case.L[:3]} n={n} N_2pi={prep.N} phi={phi:.4f} setup {setup_s:.1f}s",
          file=sys.stderr, flush=True)

    # the accuracy run: ACC_STEPS steps from zero fields (the envelope's ramp, then CW),
    # whatever --steps is (the wave crosses the grid; a few steps would compare zeros)
    acc_steps = max(ACC_STEPS, case.time_period[0] * prep.N + 1000)
    # FP32 path: the autotuner picks the step kernel (fused tile shape or the two sweeps)
    with prep.plan(fdtd.F32) as p32:
        if dft_planes:
            p32.set_dft(dft_planes, dft_w, chunk=dft_chunk)
        choice = p32.autotune(20)
        fx, fy, fz, variant = choice
        step_kernels = p32.kernels.split(" + ")
        kernel = p32.kernels + (", z chunk %d" % fz if variant else "")
        if variant == 3 and p32.z_balance:
            kernel += " (chunks balanced, z-PML stage counted %g)" % (1 + p32.z_balance / 2)
        print(f"[bench] autotuner: {kernel}", file=sys.stderr, flush=True)
        wall_ms, ev_ms = time_plan(p32, K, W, torch)
        acc32 = None if args.no_fp64 else run_from_zero(p32, dft_planes, dft_w, dft_chunk, acc_steps)
    ms_step = wall_ms / K
    gcells = n / (ms_step * 1e-3) / 1e9
    b32 = bytes_per_cell(phi, 4, 4)
    kern_ms = ev_ms / K
    achieved = b32 * n / (kern_ms * 1e-3) / 1e9
    print(f"[bench] fp32: {ms_step:.4f} ms/step, {gcells:.2f} Gcell/s, {achieved:.0f} GB/s "
          f"algorithmic", file=sys.stderr, flush=True)

    fp64 = None
    err = None
    if not args.no_fp64:
        with prep.plan(fdtd.F64) as p64:
            if dft_planes:
                p64.set_dft(dft_planes, dft_w, chunk=dft_chunk)
            wall64, ev64 = time_plan(p64, K, W, torch)
            acc64 = run_from_zero(p64, dft_planes, dft_w, dft_chunk, acc_steps)
        ms64 = wall64 / K
        g64 = n / (ms64 * 1e-3) / 1e9
        b64 = bytes_per_cell(phi, 8, 8)
        ach64 = b64 * n / (ev64 / K * 1e-3) / 1e9
        fp64 = dict(value=round(g64, 4), ms_per_step=round(ms64, 5),
                    roofline=dict(bound="hbm", achieved=round(ach64, 1), peak=HBM_PEAK_GBS,
                                  unit="GB/s", frac=round(ach64 / HBM_PEAK_GBS, 4),
                                  bytes_per_cell=round(b64, 2)),
                    speedup_fp32_over_fp64=round(gcells / g64, 3))
        err = s_comparison(prep, acc32, acc64, dft_planes)
        err.update(steps=acc_steps, ramp_steps=case.time_period[0] * prep.N, bar=1e-3,
                   passes=bool(err["S_rel_err_max"] <= 1e-3 and err["S_abs_err_max"] <= 1e-3),
                   dft=dict(planes=dft_planes, w_per_step=[round(x, 6) for x in dft_w], chunk=dft_chunk)
                   if dft_planes else None,
                   definition="|S| = |overlap with the source profile| / envelope sum of the steps run; "
                              "error = max over the reference monitors and every DFT (plane, frequency) of "
                              "| |S_fp32| - |S_fp64| |, relative = that / |S_fp64|")
        print(f"[bench] fp64: {ms64:.4f} ms/step, {g64:.2f} Gcell/s; |S| err {err['S_abs_err_max']:.2e} "
              f"abs, {err['S_rel_err_max']:.2e} rel over {err['compared']} values ({acc_steps} steps)",
              file=sys.stderr, flush=True)

    box = None
    if not args.no_box:
        box = box_run(args, torch, K, W)

    metal = None if args.no_metal else metal_run(args, torch, K, W)

    point = None if args.no_point else point_run(torch)

    batch = batch_small = None
    if not args.no_batch:
        batch = batch_run(args.batch_count, args.batch_steps, torch, args.small, ports=not args.small)
        print(f"[bench] batch of {args.batch_count}: {batch['value']:.2f} Gcell/s fused, "
              f"{batch['graph_replay']:.2f} graph replay "
              f"({batch['graph_launches']} launches), {batch['streams_mode']:.2f} streams, "
              f"{batch['one_after_another']:.2f} one after another", file=sys.stderr, flush=True)
        # small grids, where launching every step from the host is the cost
        batch_small = batch_run(args.batch_count, 4 * args.batch_steps, torch, args.small, **SMALL_BATCH)
        print(f"[bench] batch of {args.batch_count} small grids: {batch_small['value']:.2f} Gcell/s fused, "
              f"{batch_small['graph_replay']:.2f} graph replay "
              f"({batch_small['graph_launches']} launches), {batch_small['streams_mode']:.2f} streams, "
              f"{batch_small['one_after_another']:.2f} one after another", file=sys.stderr, flush=True)

    large = None
    if args.large:
        large = large_run(torch, args.large, args.large_depth, args.small)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_run(torch, args.small, args.adjoint_fd, args.adjoint_fd_seed, args.adjoint_steps)

    copy_gbs = copy_bandwidth(torch)
    print(f"[bench] device copy: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    traffic = None if args.small else measured_traffic(step_kernels, 2 if variant == 3 else 1)

    cpu = cpu_all = None
    if not args.no_cpu:
        cpu = cpu_baseline(prep, args.cpu_seconds)
        print(f"[bench] cpu baseline: {cpu['value']} Gcell/s ({cpu['kind']})", file=sys.stderr, flush=True)
        cpu_all = cpu_baseline(prep, args.cpu_seconds / 2, all_cores=True)
        if cpu_all:
            print(f"[bench] cpu baseline, {cpu_all['cores']} threads: {cpu_all['value']} Gcell/s "
                  f"({cpu_all['kind']})", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": round(gcells, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "curvilinear split-field FDTD 160x160x320: Si core "
                               "eps 12.1 in eps 2.1, scale factors 1+0.3*g (8 Fourier modes), "
                               "5% PEC edges, 10-cell z PML, PEC x/y walls, modal sheet source, "
                               "reference DFT monitors plus a 16-frequency DFT (FP64 accumulators) on 2 port planes",
                   "grid": list(case.L[:3]), "cells": n, "N_2pi": prep.N,
                   "pml_fraction": round(phi, 4), "pec_edges": case.meta.get("pec_edges"),
                   "small": bool(args.small)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "achieved_is": "algorithmic bytes (B_alg per cell x cells) / kernel time per step; the "
                                    "two-step pass reads and writes once per two steps, so this effective "
                                    "rate can exceed the device copy rate",
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_unit": "B per step (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-scaled)",
                     "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT) + " (committed counter run of "
                                       "the same kernel; null when it does not list this run's kernel)",
                     "measured_GBps": None if traffic is None else round(traffic / (kern_ms * 1e-3) / 1e9, 1),
                     "measured_frac": None if traffic is None else
                     round(traffic / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "limiter": "L1-to-L2 read requests in flight x L2 latency (TCP_TCC_READ_REQ), not "
                                "HBM bandwidth" if variant == 3 else "HBM",
                     "algorithmic_bytes_per_step": round(b32 * n),
                     "kernel": kernel + " (one step)", "copy_GBps": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4),
                     "bytes_per_cell": round(b32, 2), "kernel_ms_per_step": round(kern_ms, 5)},
        "fp64": fp64,
        "err_vs_fp64": err,
        "pml_all_faces": box,
        "metal": metal,
        "point_source": point,
        "batch": batch,
        "batch_small_grids": batch_small,
        "adjoint": adjoint,
        "large_grid": large,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
