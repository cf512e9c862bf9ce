"""Benchmark of the DEM per-step hot path on one MI355X.

Workload: 1,000,000 three-sphere clumps (3,000,000 spheres) settling in a 0.5 x 0.5 x 1.0 m plane box, FP32 state
with FP64 positions, E=5e7, nu=0.3, CoR=0.6, mu=0.4, dt=5e-6 s. Two runs: contact detection every step (headline
`value`) and every 5 steps with a drift margin of 0.2 r_min. A step = contact detection (when due) + Hertz-Mindlin
force + owner reduction + integration, all on device; state stays resident in HBM.

Prints ONE JSON line. `roofline` describes the kernel with the largest device time in the timed region:
algorithmic bytes per launch (unique compulsory traffic of the arrays it touches, actual counts of that step) over
its average HIP-event duration. `cpu_baseline` times the oracle (host C++/OpenMP restatement) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dem-engine-sbel_amd"))

HBM_PEAK_GBPS = 8000.0
_T0 = time.perf_counter()


def note(msg):
    """Progress of the legs on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def bytes_model(name, k):
    """Algorithmic bytes per launch of each kernel (counts k of the step it ran in)."""
    S, O, P, C, Np, Ba, nb, Nc = k["S"], k["O"], k["P"], k["C"], k["pairs"], k["Ba"], k["nbins"], k["cells"]
    Nk = k.get("cand", 0)
    owner_cd = 8 + 6 + 16 + 1          # voxel id, sub-voxel position, quaternion, family
    owner_dyn = 8 + 6 + 16 + 12 + 12 + 4  # + velocity, angular velocity, mass
    sph = 4 + 4 + 12 + 4               # owner, radius, offset, material
    model = {
        # sphere -> bins: templates + owner poses in, counts + CD records out, one 4-byte RMW per touched bin
        "sphere_bins": S * sph + O * owner_cd + S * (2 + 4 + 32) + P * 8,
        "bin_stats": nb * 4,
        "scan_bins": nb * 4 * 3,
        # second pass over spheres: CD records in, RMW per touch, (sphere, bin) entries out
        "bin_scatter": S * (32 + 4) + P * (4 + 8 + 8),
        # per-bin pair test: entries + bin ends + CD records (unique) + owner/family of hit pairs, pairs out
        "bin_pairs": P * (8 + 4) + S * 32 + Np * (8 + 8 + 8),
        # cell path: rows and mirror rows in one pass: permutation, two gathered counts, plane counts in; two starts out
        # candidate lists: per sphere its row and transposed-row bounds and plane count in, the two counts through the
        # tile buffer, both starts out; per candidate its flag and transposed index, its flag again through the
        # transpose; per touching sphere-sphere contact its mirror rank out
        "scan_rows": (S * (4 + 4 + 4 + 8 + 8 + 4 + 4) + Nk * (1 + 4 + 1) + Np * 4 if Nk else
                      S * (4 + 4 + 4 + 4 + 4 + 4) if k.get("cells") else S * (4 + 4 + 4)),
        "row_fill": Np * (8 + 4 + 4 + 8 + 5),
        "geo_fill": S * (4 + 4 + 32) + k["geo"] * 5,
        "row_finalize": S * 8 + C * 8,
        "history_map": S * 8 + C * 5 + k["Cprev"] * 5 + C * 4,
        "scan_owners": O * 12,
        "bside_fill": C * 5 + Np * (4 + 8 + 4),
        "bside_sort": O * 8 + Np * 8,
        # per contact: ids + type, history in/out, force record out; owner and sphere records (unique footprint)
        "force": C * (9 + 16 + 16 + 48) + S * sph + O * owner_dyn + (C * 4 if k.get("fresh") else 0),
        "anal_collect": C * 5 + k["geo"] * (4 + 48),
        # cell-list detection: per sphere template + owner pose in; cell, slot, bin count, plane count and plane bits
        # out; one 4-byte RMW per occupied cell
        "cell_assign": S * sph + O * owner_cd + S * (4 + 4 + 2 + 4 + 8) + Nc * 8,
        "scan_cells": Nc * 8,
        # template + owner pose (the CD record recomputed), cell and slot in; sorted CD record, (owner, family),
        # sorted position and the two cleared pair counts out; one cell start per sphere
        "cell_scatter": S * sph + O * owner_cd + S * (4 + 4 + 32 + 8 + 4 + 8) + Nc * 4,
        # each CD record and cell boundary once, the lane's owner/family, rowCount out; per accepted partner a slab
        # entry, its owner/family and a 4-byte RMW of the owner's B-side count
        "cell_pairs": S * (32 + 4 + 4) + O * 4 + Nc * 4 + Np * (4 + 4 + 8),
        "cell_rows": S * (4 + 8) + Np * 4 + C * 9,
        "fill": (Nc + O + 1) * 4 / 2,  # two launches per pass: cell counts and owner B-side counts
        # candidate lists: per sphere template + owner pose and the build position (fixed point) in; CD record, plane
        # count and plane bits out
        "sphere_pos": S * sph + O * owner_cd + S * 12 + S * (32 + 4 + 8),
        # per candidate its sphere and partner ids and the touching flag out; the spheres' records once
        "cand_flags": S * 32 + Nk * (4 + 4 + 1),
        # per sphere its row and transposed-row bounds, both counts out; per candidate its flag (row order) and its
        # index and flag (transposed order)
        "cand_count": S * (4 + 4 + 4 + 4) + Nk * (1 + 4 + 1),
        # per candidate its sphere and flag; per contact the row bounds, plane count, partner, previous flag and contact
        # index, transposed slot, mirror rank and the partner's mirror row start in; idA, idB, type, mirror link,
        # contact index and map out
        "cand_fill": Nk * (4 + 1) + Np * (4 + 4 + 4 + 4 + 1 + 4 + 4 + 4 + 4) + Np * (4 + 4 + 1 + 4 + 4 + 4),
        # per sphere its plane bits; per plane contact its row bound in, idA, idB, type, mirror and map out
        "cand_planes": S * 8 + k["geo"] * (4 + 4 + 4 + 1 + 4 + 4),
        # the transpose counted from the rows: per candidate its partner in, a 4-byte RMW and its place out; then
        # placed (partner, place, row start in; index out); then each transposed row sorted in place
        "cand_tcount": Nk * (4 + 4 + 4),
        "cand_tscatter": Nk * (4 + 4 + 4 + 4),
        "cand_tsort": S * 4 + Nk * (4 + 4 + 4),
        # the build sweep: as the cell sweep's pass with every candidate kept
        "cand_build": S * (32 + 4 + 4) + O * 4 + Nc * 4 + Nk * (4 + 4 + 8),
        "scan_cand": S * 4 * 2,
        # candidate rows copied (sorted) from the slab into id order with their sphere ids, build positions out
        "cand_rows": S * (4 + 4 + 32 + 12) + Nk * (4 + 4 + 4),
        "anal_final": k["anal_blocks"] * 48 * 4,
        # per owner: own record in/out, MOI, accelerations out; A rows contiguous, B lists gathered
        "owner_update": O * (64 + 64 + 16 + 24 + 8) + C * 48 + Np * (4 + 48),
    }
    return model.get(name)


def build_scene(n_clumps, update_freq, beta_frac):
    from deme_amd import scenes

    return scenes.clump_settling(n_clumps=n_clumps, update_freq=update_freq, beta_frac=beta_frac)


def run_inspectors(dev, sc, calls=20):
    """Inspector reductions over the stepped state (DEMInspector::GetValue): the clump_max_z kernel over every sphere,
    on the whole domain and on a region compiled by hipRTC; per-launch kernel time from HIP events."""
    out = {}
    for label, region in (("clump_max_z", None), ("clump_max_z_region", "return Z < 0.0f && X > -0.1f;")):
        dev.inspect("clump_max_z", region)  # compiles a region once
        dev.reset_timing()
        dev.enable_timing(True)
        for _ in range(calls):
            value = dev.inspect("clump_max_z", region)
        dev.enable_timing(False)
        t = dev.kernel_timing()
        us = 1e3 * t["inspect"][0] / max(t["inspect"][1], 1)
        algo = sc.n_spheres * 32 + sc.n_clumps * 64  # sphere records, then each owner record once
        out[label] = {"value": float(value), "inspect_us": round(us, 2),
                      "final_us": round(1e3 * t["inspect_final"][0] / max(t["inspect_final"][1], 1), 2),
                      "GBps": round(algo / (us * 1e-6) / 1e9, 1), "algorithmic_bytes": algo}
    return out


def run_device(sc, steps, warmup, update_freq, jit=False, inspect=False, prepare=None, keep_state=False):
    """Warm-up, then `steps` timed steps without kernel events, then as many with events (per-kernel durations).
    keep_state: also return the final owner state (host arrays) in info["state"]."""
    from deme_amd.abi import ANAL_TILE, DeviceSystem

    dev = DeviceSystem(sc)
    if prepare is not None:
        prepare(dev)
    if jit:
        dev.jit_build()  # force kernel specialised by hipRTC on this workload's constants
    dev.step(warmup, update_freq)
    dev.sync()
    # the timed steps, launched as in production (no events between kernels)
    t0 = time.perf_counter()
    dev.step(steps, update_freq)
    dev.sync()
    t1 = time.perf_counter()
    # the same number of steps again with kernel-level HIP events on the library's stream (per-kernel durations for
    # the roofline; the events add a gap between launches, so this pass is not the headline time)
    dev.reset_timing()
    dev.enable_timing(True)
    dev.sync()
    te0 = time.perf_counter()
    dev.step(steps, update_freq)
    dev.sync()
    te1 = time.perf_counter()
    dev.enable_timing(False)
    timing = dev.kernel_timing()
    c = dev.counts()
    skin, n_cand, builds = dev.candidate_stats()
    insp = run_inspectors(dev, sc) if inspect else None
    info = dict(S=sc.n_spheres, O=sc.n_owners, P=c.nBinSphereTouches, C=c.nContacts, Cprev=c.nPrevContacts,
                pairs=c.nContacts - c.nSphereGeoContacts, geo=c.nSphereGeoContacts, Ba=c.nActiveBins,
                nbins=sc.params.nbX * sc.params.nbY * sc.params.nbZ, cells=c.nCells, fresh=update_freq == 1,
                anal_blocks=(c.nContacts + ANAL_TILE - 1) // ANAL_TILE, cand=n_cand, skin=skin, cand_builds=builds,
                inspectors=insp,
                evented_ms_per_step=1e3 * (te1 - te0) / steps)
    if keep_state:
        info["state"] = dev.get_state()
    dev.close()
    return t1 - t0, timing, info


def run_bed(steps, warmup, settle_steps=200, cpu_threads=0, cpu_steps=2):
    """Configuration 3: 8M spheres with lognormal radii (median 1 mm, sigma 0.2, clipped to [0.5, 2] mm) and 4
    materials packed at ~0.6 solid fraction in a 0.41 x 0.41 x 0.43 m box, a top plate in a prescribed family moving
    down at 0.05 m/s, dt = 2e-6 s, contact detection every step. The packing's residual overlaps are relaxed by a
    damped run on the device first (not timed)."""
    from deme_amd import scenes

    note("polydisperse bed: packing 8M spheres")
    sc = scenes.polydisperse_bed()
    note("polydisperse bed: settling and timing on the device")
    t, timing, info = run_device(sc, steps, warmup, 1, prepare=lambda dev: scenes.settle_bed(dev, sc, settle_steps),
                                 keep_state=cpu_threads > 0)
    top = sorted(timing.items(), key=lambda kv: -kv[1][0])[:8]
    out = {"spheres": sc.n_spheres, "solid_fraction": round(sc.info["solid_fraction"], 4),
           "box_m": [round(b, 4) for b in sc.info["box"]], "settle_steps": settle_steps,
           "contacts": info["C"], "contacts_per_sphere": round(info["C"] / sc.n_spheres, 3), "steps": steps,
           "sphere_steps_per_s": sc.n_spheres * steps / t, "ms_per_step": 1e3 * t / steps,
           "kernels_us": {k: round(1e3 * v[0] / max(v[1], 1), 1) for k, v in top}}
    if cpu_threads > 0:
        dt, phases, runs = cpu_baseline(sc, cpu_steps, 1, cpu_threads, state=info["state"])
        out["cpu"] = cpu_row(sc.n_spheres * cpu_steps / dt, "sphere-steps/s", cpu_threads, runs, phases,
                             f"oracle from the device's final state, detection every step: median of 3 runs of "
                             f"{cpu_steps} steps after 1 warm-up step")
        out["gpu_over_cpu"] = out["sphere_steps_per_s"] / out["cpu"]["value"]
    return out


def cpu_row(value, unit, threads, runs, phases, sample):
    return {"value": value, "unit": unit, "cores": threads, "kind": "port", "sample": sample,
            "runs_s": [round(r, 3) for r in runs],
            "phase_s_per_step": {k: round(v, 6) for k, v in phases.items()} if phases else None}


def run_small(steps, warmup, cpu_threads=0):
    """Configuration 1: 20,000 spheres (r = 5 mm) settling in a 0.3 x 0.3 x 0.6 m box, dt = 1e-5 s, contact detection
    every step; the whole 1,000-step run is timed."""
    from deme_amd import scenes

    sc = scenes.sphere_settling()
    t, timing, info = run_device(sc, steps, warmup, 1)
    top = sorted(timing.items(), key=lambda kv: -kv[1][0])[:6]
    out = {"spheres": sc.n_spheres, "contacts": info["C"], "steps": steps,
           "sphere_steps_per_s": sc.n_spheres * steps / t, "ms_per_step": 1e3 * t / steps,
           "kernels_us": {k: round(1e3 * v[0] / max(v[1], 1), 1) for k, v in top}}
    if cpu_threads > 0:
        dt, phases, runs = cpu_baseline(sc, steps, 1, cpu_threads)
        out["cpu"] = cpu_row(sc.n_spheres * steps / dt, "sphere-steps/s", cpu_threads, runs, phases,
                             f"oracle, the full {steps}-step run from the initial state (median of 3 runs after 1 "
                             f"warm-up step)")
        out["gpu_over_cpu"] = out["sphere_steps_per_s"] / out["cpu"]["value"]
    return out


def run_batch(n_scenes, steps, warmup, update_freq, beta, pitch_scale=1.05, cpu_threads=0, cpu_steps=2):
    """Batched independent settling columns: one context stepping all scenes together, against the plain loop that
    runs each scene's steps in a context of its own, scene after scene (each context on its own stream; same world,
    same physics: each scene's trajectory is identical). pitch_scale: lattice pitch of the columns over the largest
    diameter (1.05: columns in free fall at the start; 0.75: neighbours overlap, the columns carry contacts)."""
    from deme_amd import scenes
    from deme_amd.abi import DeviceSystem

    sc = scenes.batched_columns(n_scenes=n_scenes, update_freq=update_freq, beta=beta, pitch_scale=pitch_scale)
    dev = DeviceSystem(sc)
    dev.step(warmup, update_freq)
    # autotuner: contact-update interval (with its drift margin for the free-fall speed bound of the columns) and
    # cell scale, timed on the live batch
    vmax = float(np.sqrt(2 * 9.81 * 0.3))
    update_freq, table = dev.autotune(freqs=(1, 2, 5, 10), cell_scales=(1.0, 1.25), trial_steps=10, max_vel=vmax,
                                      safety=1.2)
    f32 = np.float32
    beta = float(f32(vmax) * f32(sc.params.h) * f32(update_freq) * f32(1.2)) if update_freq > 1 else 0.0
    scale = (1.0, 1.25)[int(np.argmin(table[[1, 2, 5, 10].index(update_freq)]))]
    # second stage: candidate lists (automatic skin) or the cell sweep every pass, the sweep's mode and workgroup
    # size, at that interval
    sweep_skin, sweep_mode, sweep_block, sweep_table = dev.autotune_sweep(
        skins=(-1.0, 0.0), modes=(0, 2), blocks=(128, 256), trial_steps=2 * update_freq, update_freq=update_freq)
    dev.step(warmup, update_freq)
    dev.sync()
    t0 = time.perf_counter()
    dev.step(steps, update_freq)
    dev.sync()
    t_batch = time.perf_counter() - t0
    contacts = dev.counts().nContacts
    # the same steps as whole detection intervals replayed from captured HIP graphs: aligned to an interval start,
    # with both list parities' graphs captured before the timed intervals
    dev.set_step_graphs(True)
    align = (-dev.counts().nSteps) % update_freq
    dev.step(align + 2 * update_freq, update_freq)
    dev.sync()
    gsteps = -(-steps // update_freq) * update_freq
    s0 = dev.graph_stats()
    t0 = time.perf_counter()
    dev.step(gsteps, update_freq)
    dev.sync()
    t_graph = (time.perf_counter() - t0) * steps / gsteps
    graph_stats = tuple(b - a for a, b in zip(s0, dev.graph_stats()))
    dev.set_step_graphs(False)
    state = dev.get_state() if cpu_threads > 0 else None
    dev.close()
    parts = [DeviceSystem(p) for p in scenes.split_batch(sc, own_world=True)]
    for d in parts:
        d.set_beta(beta)
        d.set_cell_scale(scale)
        d.set_skin(sweep_skin)
        d.set_pair_sweep(sweep_mode, sweep_block)
        d.step(warmup, update_freq)
    for d in parts:
        d.sync()
    t0 = time.perf_counter()
    for d in parts:
        d.step(steps, update_freq)
    for d in parts:
        d.sync()
    t_loop = time.perf_counter() - t0
    for d in parts:
        d.close()
    return {"scenes": n_scenes, "spheres": sc.n_spheres, "pitch_scale": pitch_scale, "contacts": contacts,
            "update_freq": update_freq,
            "beta_m": beta, "cell_scale": scale, "autotune_ms_per_step": table.round(4).tolist(), "steps": steps,
            "scene_steps_per_s": n_scenes * steps / t_batch, "ms_per_step": 1e3 * t_batch / steps,
            "detection_skin_mode_block": [sweep_skin, sweep_mode, sweep_block],
            "detection_autotune_ms_per_step_skin_mode_block": sweep_table.round(4).tolist(),
            "graph_replay_ms_per_step": 1e3 * t_graph / steps,
            "graph_captures_replays_fallbacks_in_timed_steps": list(graph_stats),
            "loop_scene_steps_per_s": n_scenes * steps / t_loop, "loop_ms_per_step": 1e3 * t_loop / steps,
            "speedup_vs_loop": t_loop / t_batch,
            "cpu": (None if state is None else batch_cpu(sc, state, cpu_steps, cpu_threads))}


def batch_cpu(sc, state, steps, threads):
    dt, phases, runs = cpu_baseline(sc, steps, 1, threads, state=state)
    return cpu_row(sc.n_scenes * steps / dt, "scene-steps/s", threads, runs, phases,
                   f"oracle on the whole batch (all scenes in one system) from the device's final state, detection "
                   f"every step: median of 3 runs of {steps} steps after 1 warm-up step")


def run_adjoint(n, steps, window, warmup_windows=1, cpu_threads=0):
    """Differentiation workload: spheres dropped with seeded velocities onto a plane. Forward = the plain step loop
    with the contact list rebuilt every `window` steps; adjoint = deme_adjoint_forward (the same steps with a
    checkpoint per window) and deme_adjoint_backward (per window: recompute with per-step snapshots, then the reverse
    sweep), for the MSE of final positions against a target made with perturbed materials."""
    from deme_amd import scenes
    from deme_amd.abi import DeviceSystem, positions_f64

    sc = scenes.adjoint_drop(n=n, steps=steps, window=window)
    dev = DeviceSystem(sc)
    # target state: the same simulator with E x 1.1, CoR x 0.9, mu x 1.1
    pert = sc.materials.copy()
    pert["E"] *= 1.1
    pert["CoR"] *= 0.9
    pert["mu"] *= 1.1
    dev.upload_materials(pert)
    dev.step(steps, window)  # from the initial state of a fresh context
    s = dev.get_state()
    target = positions_f64(sc.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[:n]
    dev.upload_materials(sc.materials)
    # warm-up (buffers sized), then the timed legs, each from the initial state
    dev.set_state(sc.state)
    dev.clear_contacts()
    dev.adjoint_forward(warmup_windows * window, window)
    dev.adjoint_backward(np.zeros((n, 3)))
    dev.set_state(sc.state)
    dev.clear_contacts()
    dev.sync()
    t0 = time.perf_counter()
    dev.step(steps, window)
    dev.sync()
    t_fwd = time.perf_counter() - t0
    contacts = dev.counts().nContacts
    dev.set_state(sc.state)
    dev.clear_contacts()
    dev.sync()
    t0 = time.perf_counter()
    dev.adjoint_forward(steps, window)
    dev.sync()
    t_afwd = time.perf_counter() - t0
    s = dev.get_state()
    X = positions_f64(sc.params, s["voxelID"], s["locX"], s["locY"], s["locZ"])[:n]
    loss = float(((X - target) ** 2).sum(1).mean())
    t0 = time.perf_counter()
    gV, gM = dev.adjoint_backward(2.0 * (X - target) / n)
    t_bwd = time.perf_counter() - t0
    dev.close()
    out = {"spheres": n, "materials": len(sc.materials), "steps": steps, "window": window, "contacts": contacts,
           "forward_ms_per_step": 1e3 * t_fwd / steps, "checkpointed_forward_ms_per_step": 1e3 * t_afwd / steps,
           "backward_ms_per_step": 1e3 * t_bwd / steps, "adjoint_over_forward": t_bwd / t_fwd,
           "gradient_eval_over_forward": (t_afwd + t_bwd) / t_fwd, "loss": loss,
           "grad_E": gM[:, 0].tolist(), "grad_v0_norm": float(np.linalg.norm(gV))}
    if cpu_threads > 0:
        out["cpu"] = cpu_adjoint(sc, steps, window, 2.0 * (X - target) / n)
    return out


def cpu_adjoint(sc, steps, window, dLdX, repeats=3):
    """The FP64 host adjoint (the oracle's forward simulator and reverse sweep) on the same workload: the forward alone
    and forward + reverse, median of `repeats` each; one thread (the restatement is serial)."""
    os.environ.setdefault("OMP_WAIT_POLICY", "passive")  # (before the oracle's OpenMP runtime loads)
    import pyoracle  # noqa: E402  (the oracle is the reported CPU baseline only)

    ref = pyoracle.AdjointFP64(sc, steps, window)
    fwd, grad = [], []
    for k in range(repeats):
        t0 = time.perf_counter()
        ref.forward()
        fwd.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        ref.gradient(dLdX)
        grad.append(time.perf_counter() - t0)
        note(f"cpu adjoint run {k + 1}/{repeats}: forward {fwd[-1]:.2f} s, gradient {grad[-1]:.2f} s")
    tf, tg = sorted(fwd)[repeats // 2], sorted(grad)[repeats // 2]
    return {"forward_ms_per_step": 1e3 * tf / steps, "gradient_ms_per_step": 1e3 * tg / steps,
            "adjoint_over_forward": (tg - tf) / tf, "cores": 1, "kind": "port",
            "sample": f"FP64 host forward simulator and reverse sweep, the full {steps} steps (windows of {window}), "
                      f"median of {repeats}",
            "forward_runs_s": [round(x, 3) for x in fwd], "gradient_runs_s": [round(x, 3) for x in grad]}


def cpu_threads():
    """The job's CPU share: its affinity set, bounded by OMP_NUM_THREADS when the environment sets one."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def host_cpu():
    """nproc, the affinity set and lscpu's model, socket and core counts of the host the baseline runs on."""
    import subprocess

    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(sc, steps, update_freq, threads, repeats=3, state=None):
    """The oracle (host C++/OpenMP restatement of the reference's kernels) on `threads` threads: one warm-up step, then
    `repeats` timed runs of `steps` steps, each from the same state (the scene's, or `state`); the median run's
    seconds and its per-phase seconds (the reference's timer names)."""
    # (set before the oracle's OpenMP runtime loads) idle workers sleep instead of spinning, so the device legs timed
    # after a baseline keep the host's cores
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_WAIT_POLICY", "passive")
    import pyoracle  # noqa: E402  (the oracle is the reported CPU baseline only)

    ora = pyoracle.OracleSystem(sc, threads=threads)
    if state is not None:
        ora.set_state(state)
    ora.step(1, update_freq)  # warm-up step
    start = ora.get_state()
    runs = []
    for k in range(repeats):
        ora.set_state(start)
        ora.phase_times(reset=True)
        t0 = time.perf_counter()
        ora.step(steps, update_freq)
        runs.append((time.perf_counter() - t0, ora.phase_times()))
        note(f"cpu baseline run {k + 1}/{repeats}: {steps} steps in {runs[-1][0]:.2f} s")
    ora.close()
    runs.sort(key=lambda r: r[0])
    dt, phases = runs[len(runs) // 2]
    return dt, {k: v / steps for k, v in phases.items()}, [r[0] for r in runs]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--clumps", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-run-b", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-jit", action="store_true")
    ap.add_argument("--skip-inspect", action="store_true")
    ap.add_argument("--skip-bed", action="store_true")
    ap.add_argument("--bed-steps", type=int, default=100)
    ap.add_argument("--skip-small", action="store_true")
    ap.add_argument("--small-steps", type=int, default=1000)
    ap.add_argument("--adjoint-spheres", type=int, default=50_000)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-window", type=int, default=20)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the workload, copy a device buffer of this size once (a known-size reference "
                         "kernel for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; sharding over several GPUs is not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    if args.copy_probe_mib > 0:
        # a device-to-device copy of known size, through the HIP runtime the library itself uses (not torch's: a
        # process drives the device through one runtime)
        import ctypes as C

        from deme_amd.abi import load

        load()
        hip = C.CDLL("libamdhip64.so.7")
        n = C.c_size_t(args.copy_probe_mib << 20)
        src, dst = C.c_void_p(), C.c_void_p()
        ok = hip.hipMalloc(C.byref(src), n) == 0 and hip.hipMalloc(C.byref(dst), n) == 0
        ok = ok and hip.hipMemset(src, 1, n) == 0 and hip.hipMemcpy(dst, src, n, 3) == 0  # 3: device to device
        ok = ok and hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        if not ok:
            raise RuntimeError("copy probe failed")

    threads = 0 if args.no_cpu else cpu_threads()
    note("headline: building the 1M-clump scene")
    scA = build_scene(args.clumps, 1, 0.0)
    note("headline: timing")
    tA, timingA, infoA = run_device(scA, args.steps, args.warmup, 1, inspect=not args.skip_inspect)
    value = args.clumps * args.steps / tA

    note(f"headline done: {1e3 * tA / args.steps:.3f} ms/step")
    runB = None
    if not args.skip_run_b:
        note("every-5-steps run")
        scB = build_scene(args.clumps, 5, 0.2)
        tB, timingB, infoB = run_device(scB, args.steps, args.warmup, 5)
        totB = sum(v[0] for v in timingB.values())
        runB = {"clump_steps_per_s": args.clumps * args.steps / tB, "ms_per_step": 1e3 * tB / args.steps,
                "contacts": infoB["C"], "beta_m": scB.params.beta,
                "kernels_us_per_step": {k: round(1e3 * v[0] / args.steps, 2)
                                        for k, v in sorted(timingB.items(), key=lambda kv: -kv[1][0])},
                "device_ms_per_step": round(totB / args.steps, 4)}
        if threads:
            dt, phases, runs = cpu_baseline(scB, 10, 5, threads)
            runB["cpu"] = cpu_row(args.clumps * 10 / dt, "clump-steps/s", threads, runs, phases,
                                  "oracle, same workload (detection every 5 steps, drift margin 0.2 r_min): median "
                                  "of 3 runs of 10 steps after 1 warm-up step")

    bed = None
    if not args.skip_bed:
        bed = run_bed(args.bed_steps, 10, cpu_threads=threads)
        note("polydisperse bed done")

    small = None
    if not args.skip_small:
        note("20k spheres")
        small = run_small(args.small_steps, 20, cpu_threads=threads)
        note("20k spheres done")

    batch = batch_loaded = None
    if not args.skip_batch:
        note("batched columns")
        batch = run_batch(args.batch_scenes, args.batch_steps, 3, 1, 0.0)
        note("batched columns (contact-loaded)")
        batch_loaded = run_batch(args.batch_scenes, args.batch_steps, 3, 1, 0.0, pitch_scale=0.75,
                                 cpu_threads=threads)

    jit = None
    if not args.skip_jit:
        note("hipRTC-specialised force kernel")
        tJ, timingJ, _ = run_device(scA, args.steps, args.warmup, 1, jit=True)
        jit = {"clump_steps_per_s": args.clumps * args.steps / tJ, "ms_per_step": 1e3 * tJ / args.steps,
               "force_us": round(1e3 * timingJ["force"][0] / max(timingJ["force"][1], 1), 2),
               "force_us_shipped_kernel": round(1e3 * timingA["force"][0] / max(timingA["force"][1], 1), 2)}

    adjoint = None
    if not args.skip_adjoint:
        note("adjoint")
        adjoint = run_adjoint(args.adjoint_spheres, args.adjoint_steps, args.adjoint_window, cpu_threads=threads)
        note("adjoint done")

    # dominant kernel of the headline run
    dom = max(timingA.items(), key=lambda kv: kv[1][0])
    name, (ms_total, launches) = dom
    avg_ms = ms_total / max(launches, 1)
    bytes_per = bytes_model(name, infoA)
    achieved = (bytes_per / (avg_ms * 1e-3) / 1e9) if bytes_per else None
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a known-size copy)
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "hbm_traffic_clump_settling_1M.json")
    if os.path.exists(pmc_path):
        entry = json.load(open(pmc_path)).get("kernels", {}).get(name)
        if entry:
            traffic = entry["hbm_bytes_per_launch"]
    kernels = {k: {"avg_us": round(1e3 * v[0] / max(v[1], 1), 2), "launches": v[1],
                   "share": round(v[0] / sum(x[0] for x in timingA.values()), 4),
                   "GBps": (round(bytes_model(k, infoA) / (v[0] / max(v[1], 1) * 1e-3) / 1e9, 1)
                            if bytes_model(k, infoA) else None)}
               for k, v in sorted(timingA.items(), key=lambda kv: -kv[1][0])}

    cpu = None
    if threads:
        note("headline cpu baseline")
        dt, phases, runs = cpu_baseline(scA, args.cpu_steps, 1, threads)
        cpu = {"value": args.clumps * args.cpu_steps / dt, "unit": "clump-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle, same 1M-clump workload, contact detection every step: median of 3 runs of "
                         f"{args.cpu_steps} steps after 1 warm-up step, {threads} OpenMP threads (the job's CPU "
                         f"share), OMP_PROC_BIND=close, OMP_PLACES=cores, OMP_WAIT_POLICY=passive",
               "runs_s": [round(r, 3) for r in runs],
               "phase_s_per_step": {k: round(v, 5) for k, v in phases.items()},
               "host": host_cpu()}

    line = {
        "metric": "particle-steps/s (1M clumps)",
        "value": value,
        "unit": "clump-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tA / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 positions)",
        "data": "synthetic (seeded compacted lattice)",
        "config": {"workload": "clump_settling: 1M three-sphere clumps, 0.5x0.5x1.0 m plane box, dt=5e-6, "
                               "contact detection every step",
                   "clumps": args.clumps, "spheres": scA.n_spheres, "contacts": infoA["C"],
                   "ms_per_step_with_kernel_events": round(infoA["evented_ms_per_step"], 4),
                   "bin_sphere_touches": infoA["P"], "active_bins": infoA["Ba"],
                   "candidate_lists": {"skin_m": infoA["skin"], "candidates": infoA["cand"],
                                       "builds": infoA["cand_builds"]},
                   "run_every5_beta0.2r": runB, "jit_specialised": jit, "batched_columns": batch,
                   "batched_columns_loaded": batch_loaded,
                   "adjoint_drop": adjoint,
                   "inspectors": infoA["inspectors"],
                   "polydisperse_bed_8M": bed,
                   "sphere_settling_20k": small,
                   "kernels": kernels},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBPS) if achieved else None, "traffic": traffic,
                     "bytes_per_launch": bytes_per, "avg_launch_us": 1e3 * avg_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
