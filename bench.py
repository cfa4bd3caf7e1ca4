"""Throughput of the lunar-orbit interferometer hot path on one MI355X.

One step = the mixed-precision forward (3D visibilities with the occultation
mask) followed by its adjoint (masked back-projection) over the whole
"precision" workload: HEALPix RING nside=256 (786,432 pixels), 10 MHz,
10 satellites (45 baselines) x 20,000 snapshots = 900,000 visibilities,
sigma = 50 km offsets, orbit normal rotating about the spin axis.  Inputs are
synthetic (seeded) and resident in HBM before timing starts.

value = evaluated (visibility, pixel) pairs per second over forward + adjoint,
where evaluated pairs are the pairs the occultation mask lets through (the
exact count of m_kp = 1, from viss_sky_count_visible), counted once per pass.

Printed extras (keys of the JSON line):
  roofline      the dominant kernel (k_forward_quads) against its algorithmic flops
                and the blended vector peak of that mix, with its committed VALU
                issue fraction and HBM traffic, and the frozen per-pair rate beside it;
  fp64_gpu      the all-FP64 GPU direct sum on the same work (its own per-pixel
                kernels, and the mixed per-pixel kernels on the same layout), with the
                mixed result's error against it;
  cpu_baseline  the FP64 oracle on the host cores (OpenMP) on a sample of the rows,
                the reference's FP32 forward formulation on the host (throughput and
                error against FP64), the notebook-style NumPy forward (nside 16);
  oracle_5pct   the mixed V and density-compensated image of a fixed 5% of the rows
                against the FP64 oracle;
  snapshot_*    the snapshot-structured kernels on the same rows;
  hbm_passes    pix2ang (nside 256, 4096) and weights (1.8e6, 4e7 rows) HBM rates;
  batched       5 frequencies x 64 groups at nside 128: the one-launch batch against
                per-group launches, forward and density-compensated imaging.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--workload precision-nside256]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vis·pixel pairs/s fwd+adjoint, mixed vs FP64 GPU, rel L2 ≤1e-5, nside=256"
# MI355X published vector peaks (dense, non-matrix), TFLOP/s, and HBM3E bandwidth.
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
# Frozen flop convention per evaluated masked pair: 6 FP64 (3 FMA phase + mod-1
# reduction) + 34 FP32 (sin + cos counted 20, complex MAC 4, two mask dots 10).
# It charges a sin/cos to every (visibility, pixel) pair, so it is kept only as
# the "per-pair equivalent" rate beside the roofline of the kernels that run.
PAIR_FP64 = 6
PAIR_FP32 = 34
PAIR_T_PEAK = PAIR_FP32 / PEAK_FP32 + PAIR_FP64 / PEAK_FP64  # s per pair at peak
# Algorithmic work of the quad kernels (one sin/cos per quad of pixels
# (+-x, +-y, +-z), the ring factor once per run of same-ring tiles), per unit
# counted by viss_sky_count_work with the kernels' own occultation test; sin +
# cos count 20 FP32 flops as in the per-pair convention, an FMA 2:
#   (row, quad) with a visible pixel:   phase u x + v y (mul + FMA: 3 FP64) and its
#                                       mod-1 reduction (1 FP64), sin + cos (20 FP32)
#   (row, quad, ring) with a visible pixel: the ring's two multiply-adds (4 FP32)
#   (row, tile) with a visible pixel:   its four ring sums join the run's (4 FP32)
#   (row, run of up to 8 such tiles on one ring pair): ring phase w z and its
#                                       reduction (2 FP64), sin + cos (20 FP32), two
#                                       complex products into the row sum (12 FP32)
#   (row, rest pair) with a visible pixel: phase (5 FP64) + reduction (1), sin + cos (20),
#                                       two multiply-adds (4 FP32)
# The occultation classification and per-pixel tests are not counted: they decide
# which of these units run, and their cost shows in the VALU issue fraction.
QUAD_WORK = {"row_quads": (4, 20), "row_rings": (0, 4), "row_tiles": (0, 4), "row_runs": (2, 32),
             "row_rest_pairs": (6, 24)}
# Skies whose tiles take the recurrent phase (HEALPix rings: equal azimuth steps):
# A_{j+1} = 2 cos(d) A_j - A_{j-1}, one FMA (2 FP64) per (row, quad) from a tile's
# third quad on; the first two keep mul + FMA, charged per (row, tile) (+2 FP64).
QUAD_WORK_REC = {"row_quads": (3, 20), "row_rings": (0, 4), "row_tiles": (2, 4), "row_runs": (2, 32),
                 "row_rest_pairs": (6, 24)}


def quad_flops(work, rec=False):
    """(FP64 flops, FP32 flops) of one pass of the quad kernels over counted work."""
    table = QUAD_WORK_REC if rec else QUAD_WORK
    f64 = sum(work[k] * table[k][0] for k in table)
    f32 = sum(work[k] * table[k][1] for k in table)
    return f64, f32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="precision-nside256")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU time of the oracle sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the all-FP64 GPU comparison")
    ap.add_argument("--no-extras", action="store_true", help="skip the pix2ang / weights HBM measurements")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched one-launch comparison")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched comparison (profiling runs of that workload)")
    ap.add_argument("--batched-forward-only", action="store_true",
                    help="build the batched workload and run two batched forward launches, nothing else "
                         "(counter runs of the batch kernel: the per-group comparison serialises ~1e4 launches)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print the HBM-traffic summary of two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) and exit")
    ap.add_argument("--summarize-valu", metavar="PMC_DIR",
                    help="print the VALU issue fractions of a rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_VALU_TRANS_F32 "
                         "GRBM_GUI_ACTIVE run and exit")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first copy a buffer of this many MiB (HBM counter calibration under rocprofv3 --pmc)")
    return ap.parse_args()


def _pmc_rows(d):
    import csv
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(d, "*counter_collection.csv"))):
        rows += list(csv.DictReader(open(f)))
    return rows


def summarize_pmc(fetch_dir, write_dir, calib_mib=1024):
    """HBM bytes per dispatch of the hot kernels from FETCH_SIZE / WRITE_SIZE (KiB),
    each scaled by what the counter reads for a plain copy of calib_mib MiB."""
    fr, wr = _pmc_rows(fetch_dir), _pmc_rows(write_dir)
    kib = calib_mib * 1024.0
    copies_w = [float(r["Counter_Value"]) for r in wr if "copyBuffer" in r["Kernel_Name"]]
    copies_f = [float(r["Counter_Value"]) for r in fr if "copyBuffer" in r["Kernel_Name"]]
    fw = kib / max(copies_w)  # the calibration copy is the largest copy of the run
    ff = kib / max(copies_f)
    out = [f"calibration: {calib_mib} MiB copy reads FETCH_SIZE={max(copies_f):.1f} KiB, WRITE_SIZE={max(copies_w):.1f} KiB",
           f"scale factors: fetch x{ff:.4f}, write x{fw:.4f}"]
    # back-projection kernels: <MASK, CHUNKED> (<MASK, F64, CHUNKED> per pixel); both row-chunk variants
    keys = ("k_forward_quads<true,", "k_backproject_quads<true,", "k_forward_pairs<true>",
            "k_backproject_pairs<true,", "k_forward<true, false>",
            "k_backproject<true, false,", "k_forward_snap_pairs<true, 10>", "k_backproject_snap_pairs<true, 10>",
            "k_forward_snap<true, 10>", "k_backproject_snap<true, 10>", "k_forward_batch_quads<true,",
            "k_forward_batch_pairs<true>", "k_forward_batch<true, false>",
            "k_pix2ang", "k_pixel_dirs",
            "k_bp_prologue", "k_shell_hist", "k_row_weights")
    for key in keys:
        f = [r for r in fr if key in r["Kernel_Name"]]
        w = [r for r in wr if key in r["Kernel_Name"]]
        # one line per launch shape (same program order in both runs); the first is the workload's
        seen = set()
        for rf, rw in zip(f, w):
            if rf["Grid_Size"] in seen:
                continue
            seen.add(rf["Grid_Size"])
            fb, wb = float(rf["Counter_Value"]) * 1024 * ff, float(rw["Counter_Value"]) * 1024 * fw
            out.append(f"kernel={key.split('<')[0]} fetch_bytes={fb:.6e} write_bytes={wb:.6e} "
                       f"bytes_per_launch={fb + wb:.6e} grid={rf['Grid_Size']}")
    return "\n".join(out)


def committed_traffic(kernel, workload):
    """bytes_per_launch of `kernel` from profiles/pmc_traffic_<workload>.txt (None if absent)."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"kernel={kernel} "):
            for tok in line.split():
                if tok.startswith("bytes_per_launch="):
                    return float(tok.split("=")[1])
    return None


def committed_valu_issue(kernel, workload):
    """valu_issue_frac of `kernel` from profiles/valu_issue_<workload>.txt (None if absent)."""
    path = os.path.join(ROOT, "profiles", f"valu_issue_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"kernel={kernel} "):
            for tok in line.split():
                if tok.startswith("valu_issue_frac="):
                    return float(tok.split("=")[1])
    return None


def summarize_valu(d):
    """VALU issue fraction per kernel from a rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_VALU_TRANS_F32
    GRBM_GUI_ACTIVE run: a wavefront VALU instruction issues in 4 cycles on one of the 1024 SIMDs
    (256 CUs x 4), so frac = 4 SQ_INSTS_VALU / (1024 cycles); GRBM_GUI_ACTIVE is summed over the
    8 XCDs in the CSV.  Sin / cos issue in 8 cycles: the weighted fraction counts them twice."""
    rows = _pmc_rows(d)
    by = {}
    for r in rows:
        key = (r["Dispatch_Id"], r["Kernel_Name"], r["Grid_Size"])
        by.setdefault(key, {})[r["Counter_Name"]] = float(r["Counter_Value"])
    out = []
    seen = set()
    for (did, name, grid), c in sorted(by.items(), key=lambda kv: int(kv[0][0])):
        short = name.split("(")[0].replace("void ", "").split("<")[0].split("::")[-1]
        if (short, grid) in seen or "GRBM_GUI_ACTIVE" not in c or "SQ_INSTS_VALU" not in c:
            continue
        if not short.startswith("k_") or c["GRBM_GUI_ACTIVE"] < 8e6:  # kernels of a millisecond and more
            continue
        seen.add((short, grid))
        cyc = c["GRBM_GUI_ACTIVE"] / 8.0
        valu, trans = c["SQ_INSTS_VALU"], c.get("SQ_INSTS_VALU_TRANS_F32", 0.0)
        out.append(f"kernel={short} valu_issue_frac={4 * valu / (1024 * cyc):.4f} "
                   f"valu_issue_frac_trans8={4 * (valu + trans) / (1024 * cyc):.4f} sq_insts_valu={valu:.6e} "
                   f"sq_insts_valu_trans_f32={trans:.6e} gpu_cycles={cyc:.6e} grid={grid}")
    return "\n".join(out)


def main():
    args = parse()
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc))
        return
    if args.summarize_valu:
        print(summarize_valu(args.summarize_valu))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    from viss_optimize_3d_amd import geometry as geo
    from viss_optimize_3d_amd import ops

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calib_copy_mib > 0:
        src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = src.clone()
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calib_copy_mib} MiB done")
        del src, dst
    if args.batched_only:
        print(json.dumps({"batched": batched(ops, geo, dev)}), flush=True)
        return
    if args.batched_forward_only:
        batched_forward_launches(ops, geo, dev)
        return
    wl = geo.WORKLOADS[args.workload]
    log(f"[bench] generating {wl.name}: nside={wl.nside} K={wl.n_vis}")
    inp = geo.generate(wl)
    K = inp.uvw.shape[1]

    # pixel directions on the GPU: RING pix2ang -> (l, m, n), I = B * 4 pi / P
    th, ph = ops.pix2ang(wl.nside, nest=False, with_ring=False)
    B = torch.from_numpy(inp.B).to(dev)
    lmn, I = ops.pixel_dirs(th, ph, B)
    uvw = torch.from_numpy(inp.uvw).to(dev)
    xa = torch.from_numpy(inp.xyz_a).to(dev)
    xb = torch.from_numpy(inp.xyz_b).to(dev)
    R = geo.R_MOON
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sky = ops.Sky(lmn)
    torch.cuda.synchronize()
    sky_ms = (time.perf_counter() - t0) * 1e3
    visible = int(sky.count_visible(xa, xb, R_moon=R).item())
    dense = K * sky.P
    log(f"[bench] sky built in {sky_ms:.1f} ms; visible pairs {visible} of {dense} ({visible / dense:.3f})")

    V = torch.zeros(K, dtype=torch.complex128, device=dev)
    x = torch.zeros(sky.P, dtype=torch.float64, device=dev)
    y = torch.empty(K, dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream()

    def step(fp64=False, ev=None):
        V.zero_()
        x.zero_()
        if ev:
            ev[0].record(stream)
        sky.forward(I, uvw, xa, xb, mask=True, R_moon=R, fp64=fp64, out=V)
        if ev:
            ev[1].record(stream)
        y.copy_(V)
        if ev:
            ev[2].record(stream)
        sky.backproject(uvw, y, None, xa, xb, mask=True, R_moon=R, fp64=fp64, out=x)
        if ev:
            ev[3].record(stream)

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(ev=evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = [e[0].elapsed_time(e[1]) for e in evs]
    adj_ms = [e[2].elapsed_time(e[3]) for e in evs]
    ms_per_step = wall * 1e3 / args.steps
    # a pair is counted once per forward + adjoint step (evaluated pairs / (t_fwd + t_adj))
    pairs_per_step = visible
    value = pairs_per_step / (wall / args.steps)
    log(f"[bench] {args.steps} steps: {ms_per_step:.1f} ms/step; fwd {np.mean(fwd_ms):.1f} ms adj {np.mean(adj_ms):.1f} ms")
    V_mixed = V.clone()
    x_mixed = x.clone()

    out = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64-phase+f32-sincos", "data": "synthetic",
        "config": {"workload": wl.name, "nside": wl.nside, "npix": sky.P, "n_vis": K, "freq_hz": wl.freq_hz,
                   "n_sat": wl.n_sat, "sigma_m": wl.sigma_m, "n_snap": wl.n_snap, "mask": True,
                   "scheme": "RING", "seed": wl.seed},
    }
    fwd_avg, adj_avg = float(np.mean(fwd_ms)), float(np.mean(adj_ms))
    layout = sky.pairs()
    paired, quads = layout["uses_pairs"], layout["uses_quads"]
    kf, kb = (("k_forward_quads", "k_backproject_quads") if quads else
              ("k_forward_pairs", "k_backproject_pairs") if paired else ("k_forward", "k_backproject"))
    dom, dom_ms = (kf, fwd_avg) if fwd_avg >= adj_avg else (kb, adj_avg)
    if quads:
        work = sky.count_work(xa, xb, R_moon=R)
        assert work["visible_pairs"] == visible
        rec = layout["quad_recurrence"]
        f64, f32 = quad_flops(work, rec)
        t_peak = f32 / PEAK_FP32 + f64 / PEAK_FP64
        out["roofline"] = {
            "bound": "valu", "kernel": dom, "achieved": (f32 + f64) / (dom_ms * 1e-3) / 1e12,
            "peak": (f32 + f64) / t_peak / 1e12, "unit": "TFLOP/s", "frac": t_peak / (dom_ms * 1e-3),
            "traffic": committed_traffic(dom, wl.name), "kernel_ms": dom_ms,
            "flops_per_launch": {"fp64": f64, "fp32": f32}, "work_per_launch": work,
            "flops_per_unit": {k: {"fp64": a, "fp32": b}
                               for k, (a, b) in (QUAD_WORK_REC if rec else QUAD_WORK).items()},
            "recurrent_phase": rec,
            "note": ("algorithmic flops of the quad kernels that run (units counted exactly by "
                     "viss_sky_count_work); peak = the blended FP32 / FP64 vector peak for that mix; "
                     "frac = time at peak / kernel time"),
        }
        issue = committed_valu_issue(dom, wl.name)
        if issue is not None:
            out["roofline"]["valu_issue_frac"] = issue
    else:
        out["roofline"] = {"bound": "valu", "kernel": dom, "achieved": None, "peak": None, "unit": "TFLOP/s",
                           "frac": None, "traffic": committed_traffic(dom, wl.name),
                           "note": "the sky has no quad layout: no algorithmic count for its kernels"}
    # the frozen per-(visibility, pixel)-pair convention, as an equivalent rate
    eq = visible * (PAIR_FP32 + PAIR_FP64) / (dom_ms * 1e-3)
    out["roofline"]["per_pair_equivalent"] = {
        "achieved": eq / 1e12, "peak": (PAIR_FP32 + PAIR_FP64) / PAIR_T_PEAK / 1e12, "unit": "TFLOP/s",
        "ratio": eq * PAIR_T_PEAK / (PAIR_FP32 + PAIR_FP64), "pairs_per_launch": visible,
        "flops_per_pair": {"fp32": PAIR_FP32, "fp64": PAIR_FP64},
        "note": ("the frozen count charges a sin/cos to every visible (visibility, pixel) pair; the quad kernels "
                 "share one sin/cos among four pixels, so this ratio is not a fraction of peak and exceeds 1")}
    out["detail"] = {
        "visible_pairs_per_pass": visible, "dense_pairs_per_pass": dense, "visible_fraction": visible / dense,
        "fwd_ms": fwd_avg, "adj_ms": adj_avg, "fwd_pairs_per_s": visible / (fwd_avg * 1e-3),
        "adj_pairs_per_s": visible / (adj_avg * 1e-3), "dense_equiv_pairs_per_s": dense / (wall / args.steps),
        "sky_build_ms": sky_ms, "sky_layout": layout,
    }

    if not args.no_fp64:
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        step(fp64=True)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(fp64=True, ev=e)
        torch.cuda.synchronize()
        w64 = time.perf_counter() - t0
        v64 = pairs_per_step / w64
        rel_v = float(torch.linalg.vector_norm(V_mixed - V) / torch.linalg.vector_norm(V))
        rel_x = float(torch.linalg.vector_norm(x_mixed - x) / torch.linalg.vector_norm(x))
        max_v = float((V_mixed - V).abs().max() / V.abs().max())
        max_x = float((x_mixed - x).abs().max() / x.abs().max())
        out["fp64_gpu"] = {"value": v64, "unit": "pairs/s", "ms_per_step": w64 * 1e3,
                           "fwd_ms": e[0].elapsed_time(e[1]), "adj_ms": e[2].elapsed_time(e[3]),
                           "speedup_mixed_over_fp64": value / v64, "rel_l2_forward": rel_v,
                           "rel_l2_adjoint": rel_x, "max_err_forward_rel_to_max": max_v,
                           "max_err_adjoint_rel_to_max": max_x}
        log(f"[bench] fp64 GPU: {w64 * 1e3:.1f} ms/step; mixed/fp64 {value / v64:.2f}x; rel L2 fwd {rel_v:.2e} adj {rel_x:.2e}")
    if not args.no_fp64:
        # the precision-only comparison: the mixed per-pixel kernels (the all-FP64 path's layout) on the
        # same directions, each moved by ~1e-9 so that no pixel pairs with its antipode
        rng_p = np.random.default_rng(99)
        lmn_p = lmn + torch.from_numpy(rng_p.normal(0.0, 1e-9, lmn.shape)).to(dev)
        lmn_p = lmn_p / torch.linalg.vector_norm(lmn_p, dim=0, keepdim=True)
        sky_p = ops.Sky(lmn_p)
        assert not sky_p.pairs()["uses_pairs"]
        vis_p = int(sky_p.count_visible(xa, xb, R_moon=R).item())
        Vp = torch.zeros_like(V)
        xp = torch.zeros_like(x)

        def step_p():
            Vp.zero_()
            xp.zero_()
            sky_p.forward(I, uvw, xa, xb, mask=True, R_moon=R, out=Vp)
            sky_p.backproject(uvw, Vp, None, xa, xb, mask=True, R_moon=R, out=xp)

        step_p()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step_p()
        torch.cuda.synchronize()
        wp = time.perf_counter() - t0
        out["fp64_gpu"]["same_layout"] = {
            "mixed_per_pixel_ms_per_step": wp * 1e3, "fp64_per_pixel_ms_per_step": out["fp64_gpu"]["ms_per_step"],
            "speedup_mixed_over_fp64": out["fp64_gpu"]["ms_per_step"] / (wp * 1e3), "visible_pairs": vis_p,
            "note": "both with the per-pixel kernels (one sin/cos per visibility-pixel pair): the precision-only "
                    "speedup; speedup_mixed_over_fp64 above also has the quad kernels' shared sin/cos"}
        log(f"[bench] same layout: mixed per-pixel {wp * 1e3:.1f} ms/step, all-FP64 {w64 * 1e3:.1f} ms/step "
            f"({w64 / wp:.2f}x)")
        del sky_p, Vp, xp, lmn_p
    V64 = V.clone() if not args.no_fp64 else None  # the all-FP64 results, for the snapshot paths' errors
    x64 = x.clone() if not args.no_fp64 else None

    # snapshot-structured adjoint on the same rows (one phasor per satellite and pixel)
    pos = torch.from_numpy(inp.positions).to(dev)
    xs = torch.zeros(sky.P, dtype=torch.float64, device=dev)
    sky.backproject_snapshots(pos, wl.wavelength, V_mixed, None, mask=True, R_moon=R, out=xs)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        xs.zero_()
        sky.backproject_snapshots(pos, wl.wavelength, V_mixed, None, mask=True, R_moon=R, out=xs)
    e1.record(stream)
    torch.cuda.synchronize()
    snap_ms = e0.elapsed_time(e1) / 3
    out["snapshot_adjoint"] = {
        "ms": snap_ms, "pairs_per_s": visible / (snap_ms * 1e-3), "speedup_over_per_row_adjoint": adj_avg / snap_ms,
        "rel_l2_vs_per_row_adjoint": float(torch.linalg.vector_norm(xs - x_mixed) / torch.linalg.vector_norm(x_mixed)),
        "fwd_plus_snapshot_adjoint_pairs_per_s": visible / ((fwd_avg + snap_ms) * 1e-3)}
    if x64 is not None:
        out["snapshot_adjoint"]["rel_l2_vs_fp64_gpu"] = float(torch.linalg.vector_norm(xs - x64) /
                                                             torch.linalg.vector_norm(x64))
    log(f"[bench] snapshot adjoint: {snap_ms:.1f} ms ({adj_avg / snap_ms:.2f}x the per-row adjoint)")
    Vs = torch.zeros_like(V_mixed)
    sky.forward_snapshots(I, pos, wl.wavelength, mask=True, R_moon=R, out=Vs)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(3):
        Vs.zero_()
        sky.forward_snapshots(I, pos, wl.wavelength, mask=True, R_moon=R, out=Vs)
    e1.record(stream)
    torch.cuda.synchronize()
    sf_ms = e0.elapsed_time(e1) / 3
    out["snapshot_forward"] = {
        "ms": sf_ms, "pairs_per_s": visible / (sf_ms * 1e-3), "speedup_over_per_row_forward": fwd_avg / sf_ms,
        "rel_l2_vs_per_row_forward": float(torch.linalg.vector_norm(Vs - V_mixed) / torch.linalg.vector_norm(V_mixed)),
        "fwd_plus_adjoint_pairs_per_s": visible / ((sf_ms + snap_ms) * 1e-3)}
    if V64 is not None:
        out["snapshot_forward"]["rel_l2_vs_fp64_gpu"] = float(torch.linalg.vector_norm(Vs - V64) /
                                                             torch.linalg.vector_norm(V64))
    log(f"[bench] snapshot forward: {sf_ms:.1f} ms ({fwd_avg / sf_ms:.2f}x the per-row forward)")

    if not args.no_cpu:
        out["cpu_baseline"], out["oracle_5pct"] = cpu_baseline(args, inp, lmn, I, V_mixed, x_mixed, sky, uvw, xa,
                                                               xb, R)

    if not args.no_extras:
        out["hbm_passes"] = hbm_passes(ops, inp, wl, dev)

    if not args.no_batched:
        del sky
        torch.cuda.empty_cache()
        out["batched"] = batched(ops, geo, dev)

    print(json.dumps(out), flush=True)


def cpu_baseline(args, inp, lmn, I, V_mixed, x_mixed, sky, uvw, xa, xb, R):
    """The FP64 oracle (OpenMP on the host cores) on a fixed 5% of the visibility rows.

    The precision workload's accuracy is stated on 5% of its rows (45,000 of 900,000):
    the mixed V against the oracle's V on those rows, and the mixed density-compensated
    image of those rows (weights from all rows, the mixed V) against the oracle's image
    (same weights, the oracle's V).  The rows run in chunks (the image is a sum over
    rows); the first chunk, sized to about args.cpu_seconds of oracle time, is the timed
    throughput sample of cpu_baseline."""
    import torch

    import oracle
    from viss_optimize_3d_amd import ops

    oracle.build()
    l, m, n = (lmn[i].cpu().numpy() for i in range(3))
    Ih = I.cpu().numpy()
    K = inp.uvw.shape[1]
    rng = np.random.default_rng(123)
    n5 = max(1, K // 20)
    rows5 = np.sort(rng.choice(K, n5, replace=False))
    # density-compensation weights over every row (the shell counts need all of them)
    bll32 = inp.bll.astype(np.float32)
    nr = int(np.ceil((float(bll32.max()) - 0.25) / 0.5)) + 1
    wt = ops.weights(torch.from_numpy(bll32).to(uvw.device), uvw[2].contiguous(), nr, want_dg=False)["wt"]
    wt5 = wt[torch.from_numpy(rows5).to(uvw.device)].cpu().numpy()
    # first chunk: grow a probe until its forward takes about a quarter of the sample budget
    probe = 256
    while True:
        r = rows5[:probe]
        t0 = time.perf_counter()
        oracle.forward(l, m, n, Ih, inp.uvw[:, r], inp.xyz_a[:, r], inp.xyz_b[:, r], flags=oracle.MASK, R_moon=R)
        tp = time.perf_counter() - t0
        if tp >= args.cpu_seconds / 16 or probe >= n5:
            break
        probe = min(n5, probe * 4)
    first = int(min(n5, max(probe, probe * (args.cpu_seconds / 2) / max(tp, 1e-3))))
    log(f"[bench] oracle 5%: {n5} rows, first chunk {first} rows (probe {probe} rows forward in {tp:.2f} s)")
    Vo = np.empty(n5, np.complex128)
    img_o = np.zeros(l.size)
    t_fwd = t_adj = 0.0
    sample = None
    bounds = [0, first] + list(range(first + 10_000, n5, 10_000)) + [n5]
    bounds = sorted(set(b for b in bounds if b <= n5))
    for c0, c1 in zip(bounds[:-1], bounds[1:]):
        r = rows5[c0:c1]
        u, a, b = inp.uvw[:, r], inp.xyz_a[:, r], inp.xyz_b[:, r]
        t0 = time.perf_counter()
        Vo[c0:c1] = oracle.forward(l, m, n, Ih, u, a, b, flags=oracle.MASK, R_moon=R)
        t1 = time.perf_counter()
        img_o += oracle.backproject(l, m, n, u, Vo[c0:c1], wt5[c0:c1], a, b, flags=oracle.MASK, R_moon=R)
        t2 = time.perf_counter()
        t_fwd += t1 - t0
        t_adj += t2 - t1
        if sample is None:
            sample = (c1 - c0, t1 - t0, t2 - t1)
        log(f"[bench] oracle 5%: rows {c1}/{n5} done ({t_fwd + t_adj:.1f} s)")
    rr = torch.from_numpy(rows5).to(uvw.device)
    Vg = V_mixed[rr]
    Vgh = Vg.cpu().numpy()
    sub = lambda t: t[:, rr].contiguous()  # noqa: E731
    img_g = sky.backproject(sub(uvw), Vg, wt[rr].contiguous(), sub(xa), sub(xb), mask=True, R_moon=R).cpu().numpy()
    acc = {"rows": int(n5), "fraction_of_rows": n5 / K,
           "rel_l2_forward": float(np.linalg.norm(Vgh - Vo) / np.linalg.norm(Vo)),
           "max_err_forward_rel_to_max": float(np.max(np.abs(Vgh - Vo)) / np.max(np.abs(Vo))),
           "rel_l2_image": float(np.linalg.norm(img_g - img_o) / np.linalg.norm(img_o)),
           "max_err_image_rel_to_max": float(np.max(np.abs(img_g - img_o)) / np.max(np.abs(img_o))),
           "image": "density-compensated back-projection of the 5% rows (weights over all rows): mixed V and "
                    "kernels against the oracle's V and back-projection", "oracle_seconds": t_fwd + t_adj}
    log(f"[bench] oracle 5%: V rel L2 {acc['rel_l2_forward']:.2e} (max {acc['max_err_forward_rel_to_max']:.2e}), "
        f"image rel L2 {acc['rel_l2_image']:.2e} (max {acc['max_err_image_rel_to_max']:.2e})")
    n_s, tf, ta = sample
    rs = torch.from_numpy(rows5[:n_s]).to(uvw.device)
    vis_s = int(sky.count_visible(xa[:, rs].contiguous(), xb[:, rs].contiguous(), R_moon=R).item())
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rows = rows5[:n_s]
    n_rows = n_s
    # the reference's own FP32 forward (acos mask at the fixed cut angle, phase
    # against n - 1; vissgen3d_noblockage.cu:76-102) on a quarter of the rows,
    # and its error against the FP64 oracle of the same formulation
    r4 = rows[:max(1, n_rows // 4)]
    u4, a4, b4 = inp.uvw[:, r4], inp.xyz_a[:, r4], inp.xyz_b[:, r4]
    t0 = time.perf_counter()
    V32 = oracle.forward_ref32(l, m, n, Ih, u4, a4, b4, oracle.reference_cut_angle())
    t32 = time.perf_counter() - t0
    V64 = oracle.forward(l, m, n, Ih, u4, a4, b4, flags=oracle.MASK | oracle.PHASE_TRACK, R_moon=1737.1e3,
                         orbit_radius=1737.1e3 + 300e3)
    ref32 = {"dense_pairs_per_s": r4.size * l.size / t32, "rows": int(r4.size), "seconds": t32,
             "rel_l2_vs_fp64_same_formulation": float(np.linalg.norm(V32 - V64) / np.linalg.norm(V64))}
    log(f"[bench] cpu reference FP32 formulation: {r4.size} rows in {t32:.2f} s, "
        f"rel L2 vs FP64 {ref32['rel_l2_vs_fp64_same_formulation']:.2e}")
    # the notebook-style NumPy path (one vectorised exp over the (rows x pixels) phase matrix) on the
    # CPU-runnable case: plain-nside16, 3,000 rows x 3,072 pixels, no occultation
    from viss_optimize_3d_amd import geometry as geo

    w1 = geo.WORKLOADS["plain-nside16"]
    in1 = geo.generate(w1)
    th1, ph1, _, _ = oracle.pix2ang(w1.nside, False)
    l1, m1, n1, I1 = oracle.pixel_dirs(th1, ph1, in1.B, 4.0 * np.pi / th1.size)
    t0 = time.perf_counter()
    Vn = np.exp(-2j * np.pi * (in1.uvw.T @ np.stack([l1, m1, n1]))) @ I1
    tn = time.perf_counter() - t0
    Vo1 = oracle.forward(l1, m1, n1, I1, in1.uvw)
    numpy1 = {"pairs_per_s": in1.uvw.shape[1] * l1.size / tn, "seconds": tn, "rows": int(in1.uvw.shape[1]),
              "pixels": int(l1.size), "rel_l2_vs_oracle": float(np.linalg.norm(Vn - Vo1) / np.linalg.norm(Vo1))}
    log(f"[bench] cpu NumPy notebook-style forward, {w1.name}: {tn:.2f} s, {numpy1['pairs_per_s']:.3g} pairs/s")
    base = {"value": vis_s / (tf + ta), "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"{n_rows} of {K} visibility rows x all {l.size} pixels, FP64 oracle forward + "
                      f"density-compensated back-projection (OpenMP), {tf + ta:.1f} s",
            "reference_fp32_forward": ref32, "numpy_notebook_forward": numpy1}
    return base, acc


def batched_forward_launches(ops, geo, dev, n=2):
    """The batched workload's work-table forward, n launches (default table), for counter runs."""
    import torch

    B, uvw, xa, xb, offs, gf = geo.batched_workload()
    th, ph = ops.pix2ang(128, nest=False, with_ring=False)
    lmn, I = ops.pixel_dirs(th, ph, torch.from_numpy(B).to(dev))
    sky = ops.Sky(lmn)
    U, A, Bx = (torch.from_numpy(a).to(dev) for a in (uvw, xa, xb))
    V = torch.zeros(U.shape[1], dtype=torch.complex128, device=dev)
    batch = ops.Batch(sky, offs, U, A, Bx, mask=True, R_moon=geo.R_MOON)
    log(f"[bench] batched forward only: {len(offs) - 1} groups, {U.shape[1]} rows")
    for i in range(n):
        batch.forward(I, out=V)
        torch.cuda.synchronize()
        log(f"[bench] batched forward launch {i + 1}/{n} done")


def batched(ops, geo, dev):
    """5 frequencies x 64 snapshot groups x 500 snapshots x 28 baselines at nside 128: one
    (hipGraph-replayed, autotuned) work-table launch against one launch per group."""
    import torch

    B, uvw, xa, xb, offs, gf = geo.batched_workload()
    th, ph = ops.pix2ang(128, nest=False, with_ring=False)
    lmn, I = ops.pixel_dirs(th, ph, torch.from_numpy(B).to(dev))
    sky = ops.Sky(lmn)
    U, A, Bx = (torch.from_numpy(a).to(dev) for a in (uvw, xa, xb))
    R = geo.R_MOON
    visible = int(sky.count_visible(A, Bx, R_moon=R).item())
    G = len(offs) - 1
    groups = [(U[:, offs[g]:offs[g + 1]].contiguous(), A[:, offs[g]:offs[g + 1]].contiguous(),
               Bx[:, offs[g]:offs[g + 1]].contiguous()) for g in range(G)]
    V = torch.zeros(U.shape[1], dtype=torch.complex128, device=dev)
    log(f"[bench] batched: {G} groups, {U.shape[1]} rows, visible pairs {visible}")

    def per_group():
        for g, (u, a, b) in enumerate(groups):
            sky.forward(I, u, a, b, mask=True, R_moon=R, out=V[offs[g]:offs[g + 1]])

    batch = ops.Batch(sky, offs, U, A, Bx, mask=True, R_moon=R)
    st0 = batch.stats()
    side = torch.cuda.Stream()

    def timed(fn, label, reps=3):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        log(f"[bench] batched: {label} timed ({t * 1e3:.1f} ms)")  # progress (a counter run serialises ~1e4 launches)
        return t

    t_group = timed(per_group, "per-group forward")
    V.zero_()
    per_group()
    V_group = V.clone()
    t_default = timed(lambda: batch.forward(I, out=V), "batch forward")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.autotune(I, reps=1)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        t_graph = timed(lambda: batch.forward(I, out=V, use_graph=True), "tuned batch forward (graph)")
        V.zero_()
        batch.forward(I, out=V, use_graph=True)
    torch.cuda.synchronize()
    rel = float(torch.linalg.vector_norm(V - V_group) / torch.linalg.vector_norm(V_group))
    st = batch.stats()
    log(f"[bench] batched: graph-replayed tuned batch vs per-group launches rel L2 {rel:.2e}")
    # density-compensated imaging of every group: per-group weights (one nr for all groups) and
    # the summed image, batched (two weight launches, one back-projection graph) against per-group
    # viss_weights + viss_sky_backproject launches
    bll = torch.sqrt((U * U).sum(0)).float()
    nr = int(torch.ceil((bll.max() - 0.25) / 0.5).item()) + 1
    yv = V_group
    x_b = torch.zeros(sky.P, dtype=torch.float64, device=dev)
    x_g = torch.zeros_like(x_b)
    wt_g = torch.empty(U.shape[1], dtype=torch.float64, device=dev)

    def img_group():
        x_g.zero_()
        for g, (u, a, b) in enumerate(groups):
            s = slice(offs[g], offs[g + 1])
            wt_g[s] = ops.weights(bll[s], u[2], nr, want_dg=False)["wt"]
            sky.backproject(u, yv[s], wt_g[s], a, b, mask=True, R_moon=R, out=x_g)

    def img_batch():
        w = batch.weights(bll, nr, want_dg=False)["wt"]
        x_b.zero_()
        batch.backproject(yv, w, out=x_b, use_graph=True)

    t_img_group = timed(img_group, "per-group imaging")
    with torch.cuda.stream(side):
        side.wait_stream(torch.cuda.current_stream())
        t_img_batch = timed(img_batch, "batch imaging")
    torch.cuda.synchronize()
    rel_img = float(torch.linalg.vector_norm(x_b - x_g) / torch.linalg.vector_norm(x_g))
    log(f"[bench] batched imaging: {t_img_batch * 1e3:.1f} ms vs per-group {t_img_group * 1e3:.1f} ms, "
        f"rel L2 {rel_img:.2e}")
    imaging = {"batched_ms": t_img_batch * 1e3, "per_group_ms": t_img_group * 1e3,
               "batched_pairs_per_s": visible / t_img_batch, "per_group_pairs_per_s": visible / t_img_group,
               "speedup_batched_over_per_group": t_img_group / t_img_batch, "rel_l2_batched_vs_per_group": rel_img,
               "nr": nr, "what": "per-group density-compensation weights and the summed back-projected image"}
    return {"groups": G, "rows": int(U.shape[1]), "npix": sky.P, "visible_pairs": visible,
            "visible_tile_fraction": st["visible_tiles"] / st["total_tiles"],
            "per_group_launches_pairs_per_s": visible / t_group, "batched_default_pairs_per_s": visible / t_default,
            "batched_tuned_graph_pairs_per_s": visible / t_graph, "speedup_batched_over_per_group": t_group / t_graph,
            "rel_l2_batched_vs_per_group": rel,
            "autotune_s": t_tune, "work_items": st["n_items"], "work_items_untuned": st0["n_items"],
            "imaging": imaging}


def hbm_passes(ops, inp, wl, dev):
    """HBM rates of the O(P) / O(K) passes (algorithmic bytes / GPU time).  Each pass is
    captured 20 times into a HIP graph and replayed, so the host's per-call launch cost
    (larger than the 10-us kernels of the small sizes) stays out of the time.  The calls
    go straight to the C ABI on preallocated buffers (no allocation or memset is captured)."""
    import torch

    from viss_optimize_3d_amd import _lib

    L = _lib.load()
    res = {}

    def timed(fn, reps=20):
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(3):
                fn()  # warm (and the first-use costs of the library)
        side.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side, capture_error_mode="relaxed"):
            for _ in range(reps):
                fn()
        g.replay()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(3):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            g.replay()
            b.record()
            torch.cuda.synchronize()
            best = min(best, a.elapsed_time(b) / reps * 1e-3)
        del g
        return best

    for nside in (256, 4096):
        P = 12 * nside * nside
        th = torch.empty(P, dtype=torch.float64, device=dev)
        ph = torch.empty(P, dtype=torch.float64, device=dev)

        def f():
            ops.check("viss_pix2ang", L.viss_pix2ang(nside, _lib.SCHEME_RING, ops._p(th), ops._p(ph), None, None,
                                                     ops._stream()))

        t = timed(f)
        res[f"pix2ang_ring_nside{nside}"] = {"bytes": 16 * P, "s": t, "GB_per_s": 16 * P / t / 1e9,
                                             "frac_hbm": 16 * P / t / PEAK_HBM}
        del th, ph
    from viss_optimize_3d_amd import geometry as geo

    for rows in (2 * inp.uvw.shape[1], 40_000_000):
        rng = np.random.default_rng(1)
        bll = torch.from_numpy(rng.uniform(0, 1.2 * geo.REF_BL_MAX / wl.wavelength * 2, rows).astype(np.float32)).to(dev) \
            if rows != 2 * inp.uvw.shape[1] else torch.from_numpy(np.concatenate([inp.bll, inp.bll]).astype(np.float32)).to(dev)
        w = (torch.rand(rows, dtype=torch.float64, device=dev) * 2 - 1) * bll.double() * 0.9
        nr = int(math.ceil(float(bll.max()) * 2))  # cover every shell (no out-of-range rows)
        gs = torch.empty(rows, dtype=torch.int32, device=dev)
        mb = torch.zeros(nr, dtype=torch.int32, device=dev)  # accumulates across calls: timing only
        dcf = torch.empty(nr + 1, dtype=torch.float64, device=dev)
        wt = torch.empty(rows, dtype=torch.float64, device=dev)

        def fw():
            ops.check("viss_weights", L.viss_weights(ops._p(bll), ops._p(w), rows, nr, 0, ops._p(gs), ops._p(mb),
                                                     ops._p(dcf), None, ops._p(wt), None, ops._stream()))

        t = timed(fw)
        # algorithmic bytes per row (the two-pass convention of the weights: the histogram
        # pass reads bll, the weight pass reads bll and w and writes gs and wt, since no
        # weight is known before every row is counted): 4 + 4 + 8 (w, FP64) + 4 + 8 (wt, FP64)
        byt = 28 * rows
        res[f"weights_rows{rows}"] = {"nr": nr, "bytes": byt, "s": t, "GB_per_s": byt / t / 1e9,
                                      "frac_hbm": byt / t / PEAK_HBM}
        del bll, w, gs, mb, dcf, wt
    return res


if __name__ == "__main__":
    main()
