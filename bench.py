"""Benchmark: batched forward+backward spectra/s of the equirectangular RF
rasterizer + fused-SSIM loss on one MI355X.

Workload (one "step"): 50,000 Gaussians (70% within 0.3 m of the room walls),
64 receivers, 90x360 equirect maps; per-Gaussian emission MLPs (3->32->2);
loss 0.8*L1 + 0.2*(1 - SSIM) (11x11 window, sigma 1.5) against target spectra
rendered from a second seeded scene; full backward to means, scales,
rotations, opacities and MLP weights.  All inputs resident on the GPU.
The timed step is the library's HIP graph of forward (tile lists sized by a
capacity), fused loss and backward, replayed; the eager torch-autograd form
of the same step is timed and reported beside it (--eager times only that).

Prints ONE JSON line.  Also reported: the same step done the plain way (one
receiver per call, the reference's calling pattern), p50/p99 single-spectrum
forward latency, per-kernel times from HIP events, and the CPU oracle timed
on a bounded sample as a baseline.

--workload scenes measures the uneven-scene batch instead: 16 independent
scenes with Gaussian counts log-uniform in [5k, 200k] and 4..64 receivers
each, forward+backward of all of them in one launch sequence; it adds p50/p99
of the batch step and of single spectra rendered alone.

--workload config1 measures the single-receiver training step with the shared
emission MLP (20k Gaussians) and, beside it, the MLP's MFMA kernels at 3.2M
(receiver, Gaussian) rows against the FP32 matrix peak.

--workload csi measures multi-frequency maps: 50k Gaussians, 32 receivers,
52 subcarriers (complex, 90x360x104 floats per receiver), L1 on the complex
map + SSIM on the magnitudes, forward+backward.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gsparc-wirelessgaussiansplatting_amd")
sys.path.insert(0, PKG)

METRIC = "fwd+bwd spectra/sec, 50k Gaussians, 90×360, 64 rx batched vs plain; HBM GB/s"
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
# Algorithmic FP32 work of the blend, per (pixel, list entry) pair of the
# reference's own loops (forward.cu:198-236, backward.cu:547-615): every pair
# it visits costs the offset, power, exponential and alpha test (15 ops);
# a contributing pair's backward adds the transmittance recovery and the
# gradient terms (25 more, 40 in all).  The kernels' counting variants count
# these pairs exactly, whatever this build culls.
FWD_FLOP_PER_PAIR = 15      # per pair the reference's forward visits
BWD_FLOP_PER_PAIR = 15      # per pair the reference's backward walks
BWD_FLOP_PER_CONTRIB = 25   # more per contributing pair (backward)
CSI_FWD_FLOP_PER_CONTRIB = 2 * 2 * 52          # multi-frequency forward: 104 FMAs per contributing pair
CSI_BWD_FLOP_PER_CONTRIB = 25 + 2 * 6 * 52     # multi-frequency adjoint: 6 FMAs per subcarrier and pair
FLOP_PER_PAIR_TEXT = {
    "render_fwd": f"{FWD_FLOP_PER_PAIR} per visited pair",
    "render_bwd": f"{BWD_FLOP_PER_PAIR} per walked pair + {BWD_FLOP_PER_CONTRIB} per contributing pair",
    "render_fwd_csi": f"{FWD_FLOP_PER_PAIR} per visited pair + {CSI_FWD_FLOP_PER_CONTRIB} per contributing pair",
    "render_bwd_csi": f"{BWD_FLOP_PER_PAIR} per walked pair + {CSI_BWD_FLOP_PER_CONTRIB} per contributing pair",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--gaussians", type=int, default=50000)
    ap.add_argument("--receivers", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--latency-iters", type=int, default=100)
    ap.add_argument("--cpu-receivers", type=int, default=64,
                    help="spectra in the multi-core CPU baseline sample (0: skip the CPU baseline)")
    ap.add_argument("--cpu-receivers-1", type=int, default=16,
                    help="spectra in the single-core CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the multi-core CPU baseline (0: OMP_NUM_THREADS, at most 16)")
    ap.add_argument("--grad-mode", default="reference", choices=["reference", "exact"])
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a plain device-to-device copy of this size (counter calibration)")
    ap.add_argument("--fwd-ppl", type=int, default=0, help="forward blend pixels per lane (0: library default)")
    ap.add_argument("--bwd-ppl", type=int, default=0, help="backward blend pixels per lane (0: automatic)")
    ap.add_argument("--order", type=int, default=0, help="blend work order (0: automatic)")
    ap.add_argument("--fwd-depth", type=int, default=0, help="forward blend LDS batch depth (0: automatic)")
    ap.add_argument("--bwd-depth", type=int, default=0, help="backward blend LDS batch depth (0: automatic)")
    ap.add_argument("--workload", default="receivers", choices=["receivers", "scenes", "csi", "config1"],
                    help="receivers: one scene, R receivers (default); scenes: 16 uneven scenes in one batch; "
                         "csi: 52-subcarrier complex maps; config1: single-receiver training step with the "
                         "shared 68->128->128->1 emission MLP on MFMA")
    ap.add_argument("--csi-receivers", type=int, default=32)
    ap.add_argument("--no-tune", action="store_true", help="skip the blend-mapping autotuner")
    ap.add_argument("--eager", action="store_true",
                    help="time the eager autograd step instead of the HIP-graph-captured one")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


TRAFFIC_FILE = "profiles/traffic_50k_64rx.txt"


def measured_traffic(args, P, R):
    """HBM bytes per launch of each kernel, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised in TRAFFIC_FILE (calibrated against a copy of
    known size there).  Only for the configuration those runs measured."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if (P, R, args.grad_mode) != (50000, 64, "reference") or not os.path.exists(path):
        return {}
    out = {}
    for line in open(path):
        parts = line.split()
        if len(parts) == 4 and parts[3].startswith("traffic_bytes="):
            out[parts[0]] = int(parts[3].split("=")[1])
    return out


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


class Workload:
    """What one timed step does and how to describe it."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def loss_fn(ore, tgt):
    """0.8 * L1 + 0.2 * (1 - SSIM) on the real channel (train.py:418-427)."""
    from fused_ssim import fused_ssim
    l1 = (ore - tgt).abs().mean()
    ss = fused_ssim(ore.unsqueeze(1), tgt.unsqueeze(1))
    return 0.8 * l1 + 0.2 * (1.0 - ss)


def plain_step(params, settings_for, targets, rx_count):
    """One receiver per call through the reference API (renderer.py:39-88)."""
    import torch
    import cuda_rasterizer as cr
    for prm, st, tgt in zip(params, settings_for, targets):
        m, s_, q, o, w = prm
        m2d = torch.zeros_like(m, requires_grad=True)
        ore, oim, radii = cr.rasterize_gaussians(m, m2d, w, o, s_, q, st)
        loss = loss_fn(ore, tgt) / rx_count
        loss.backward()


def cpu_threads(args):
    """Threads of the multi-core CPU baseline: --cpu-threads, else the OpenMP
    default (OMP_NUM_THREADS), at most 16 (one GPU's share of the host)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = args.cpu_threads if args.cpu_threads > 0 else O.set_threads(0)
    return max(1, min(int(n), 16))


def cpu_baseline(args, run, n_multi, n_single, label):
    """The CPU oracle timed on a bounded sample: run(n) processes the first n
    spectra of the sample.  Multi-core over n_multi spectra with
    cpu_threads(args) OpenMP threads, then one core over n_single spectra."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nt = cpu_threads(args)
    res = {}
    for threads, n in ((nt, n_multi), (1, n_single)):
        if n <= 0:
            continue
        O.set_threads(threads)
        log(args, f"CPU baseline: oracle fwd+bwd on {n} spectra, {threads} thread(s)")
        t0 = time.perf_counter()
        run(n)
        res[threads] = (n, time.perf_counter() - t0)
    O.set_threads(nt)
    n, el = res[nt]
    out = {"value": round(n / el, 4), "unit": "spectra/s", "cores": nt, "kind": "port",
           "sample": f"{label}: first {n} spectra, fwd+bwd incl. loss, FP32 C++ oracle, {nt} OpenMP threads "
                     f"({el:.1f} s)"}
    if 1 in res and nt != 1:
        n1, el1 = res[1]
        out["single_core"] = {"value": round(n1 / el1, 4), "cores": 1,
                              "sample": f"first {n1} spectra of the same sample, one thread ({el1:.1f} s)"}
    return out


def cpu_oracle_sample(args, scenes_and_rx, targets, label):
    """The CPU oracle on a bounded sample of the same spectra: forward, loss
    gradient and backward of each sampled spectrum (multi-core, then one
    core on a smaller part of the sample)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    def run(n):
        accs = {}
        for (sc, r), tgt in list(zip(scenes_and_rx, targets))[:n]:
            acc = accs.setdefault(id(sc), O.Grads(sc.P))
            pre, keys, vals, ranges, fwd = O.render(sc, r)
            img1 = fwd["out_re"][None, None].astype(np.float32)
            img2 = tgt[None, None]
            m_, d1, d2, d3 = O.ssim_fwd(img1, img2, dtype=np.float32)
            gss = O.ssim_bwd(img1, img2, np.full_like(img1, -0.2 / img1.size), d1, d2, d3, dtype=np.float32)
            dre = (0.8 / img1.size) * np.sign(img1 - img2)[0, 0] + gss[0, 0]
            O.backward(sc, pre, vals, ranges, fwd, dre, np.zeros_like(dre), O.GRAD_REFERENCE, r=r, grads=acc)

    n = len(scenes_and_rx)
    return cpu_baseline(args, run, n, min(n, args.cpu_receivers_1), label)


def setup_receivers(args, dev):
    """One scene (config2 generator), R receivers in one batch."""
    import numpy as np
    import torch
    import cuda_rasterizer as cr
    import gsparc_scenes as S
    import wireless_rasterizer as wr
    P, R = args.gaussians, args.receivers
    sc = S.config2(P=P, R=R)
    tgt_sc = S.config2(seed=3, P=P, R=R)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    means, scales, rots, opac, mlp = (t(sc.means), t(sc.scales), t(sc.rotations), t(sc.opacities), t(sc.mlp))
    views, txp, rxp = t(sc.views), t(sc.tx_pos), t(sc.rx_pos)
    with torch.no_grad():
        target = wr.forward_batched(t(tgt_sc.means), t(tgt_sc.mlp), t(tgt_sc.opacities), t(tgt_sc.scales),
                                    t(tgt_sc.rotations), sc.H, sc.W, 1.0, views, txp, rxp)[1].clone()
    params = [x.clone().requires_grad_(True) for x in (means, scales, rots, opac, mlp)]
    settings = cr.BatchedSetting(sc.H, sc.W, 1.0, views, txp, rxp, args.grad_mode)

    def step():
        m, s_, q, o, w = params
        ore, oim, radii = cr.rasterize_gaussians_batched(m, w, o, s_, q, settings)
        loss_fn(ore, target).backward()
        for p_ in params:
            p_.grad = None

    plain_settings = [cr.GaussianRasterizerSetting(sc.H, sc.W, 1.0, views[r], torch.eye(4, device=dev), txp,
                                                   rxp[r], args.grad_mode) for r in range(R)]

    def step_plain():
        plain_step([params] * R, plain_settings, [target[r:r + 1] for r in range(R)], R)
        for p_ in params:
            p_.grad = None

    def single(i):
        wr.rasterize_gaussians(means, mlp, opac, scales, rots, sc.H, sc.W, 1.0, views[i % R], None, txp, rxp[i % R])

    def count():
        return wr.forward_batched(means, mlp, opac, scales, rots, sc.H, sc.W, 1.0, views, txp, rxp)[0]

    def tune():
        with torch.no_grad():
            out = wr.forward_batched(means, mlp, opac, scales, rots, sc.H, sc.W, 1.0, views, txp, rxp)
        return wr.autotune(out, grad_mode=args.grad_mode)

    def graphed(deterministic=False):
        """The same step (forward, fused loss, backward) captured in a HIP
        graph; tile lists sized for 1.25 K."""
        import gsparc_train as T
        cap = int(count() * 1.25) + 1024
        live = [x.clone() for x in (means, scales, rots, opac, mlp)]
        g = T.GraphedTrainingStep(*live, views, txp, rxp, target, sc.H, sc.W, capacity=cap,
                                  grad_mode=args.grad_mode, deterministic=deterministic)
        return g, live

    def cpu():
        nr = min(args.cpu_receivers, R)
        tg = target[:nr].cpu().numpy()
        return cpu_oracle_sample(args, [(sc, r) for r in range(nr)], list(tg),
                                 f"{nr} of the {R} receivers of the same {P // 1000}k-Gaussian scene")

    def trainer(use_graphs=True):
        """The whole training iteration (train.py:364-566 without schedule
        events): step, densification statistics, Adam over every group with
        the activations, centre clamp; confidence-weighted loss."""
        import gsparc_gaussian as G
        import gsparc_train as T
        import gsparc_trainer as TR
        m = G.GaussianModel()
        m.load_raw(means, torch.log(scales), rots, torch.logit(opac.double()).float(), mlp)
        torch.manual_seed(0)
        conf = T.ConfidenceMLP(32, device=dev).packed().detach().clone().contiguous()
        o = G.OptimizationParams(densify_from_iter=10 ** 9)
        return TR.Trainer(m, o, views, txp, rxp, target, sc.H, sc.W, xyz_min=[0.0, 0.0, 0.0],
                          xyz_max=list(S.ROOM), confidence=conf, grad_mode=args.grad_mode,
                          k_capacity=int(count() * 1.25) + 1024, use_graphs=use_graphs)

    return Workload(
        step=step, step_plain=step_plain, single=single, count=count, cpu=cpu, graphed=graphed, tune=tune,
        trainer=trainer,
        spectra=R, Ptot=P,
        Nrec=P * R,
        H=sc.H, W=sc.W, traffic_key=(P, R),
        data="synthetic (seeded: 70% wall-clustered Gaussians, receivers uniform in a 6x6x3 m room)",
        config={"workload": "batched receivers, fwd+bwd, 0.8*L1 + 0.2*(1-SSIM)", "gaussians": P, "receivers": R,
                "map": f"{sc.H}x{sc.W}", "emission": "per-Gaussian MLP 3->32->2", "grad_mode": args.grad_mode})


def setup_scenes(args, dev):
    """16 uneven scenes (config4 generator) in one launch sequence."""
    import numpy as np
    import torch
    import cuda_rasterizer as cr
    import gsparc_scenes as S
    import wireless_rasterizer as wr
    scs = S.config4()
    tgts = S.config4(variant=1)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    per = [dict(means=t(sc.means), scales=t(sc.scales), rots=t(sc.rotations), opac=t(sc.opacities),
                mlp=t(sc.mlp), views=t(sc.views), tx=t(sc.tx_pos), rx=t(sc.rx_pos)) for sc in scs]
    cat = lambda k: torch.cat([d[k] for d in per])
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], dev)
    views, rxp, txp = cat("views"), cat("rx"), torch.stack([d["tx"] for d in per])
    H, W = scs[0].H, scs[0].W
    with torch.no_grad():
        tm = torch.cat([t(x.means) for x in tgts])
        tmlp = wr.pack_mlps([t(x.mlp) for x in tgts])
        target = wr.forward_scenes(tm, tmlp, torch.cat([t(x.opacities) for x in tgts]),
                                   torch.cat([t(x.scales) for x in tgts]), torch.cat([t(x.rotations) for x in tgts]),
                                   H, W, 1.0, views, txp, rxp, batch)[1].clone()
        del tm, tmlp
    packed = [cat("means"), cat("scales"), cat("rots"), cat("opac"), wr.pack_mlps([d["mlp"] for d in per])]
    params = [x.clone().requires_grad_(True) for x in packed]
    settings = cr.SceneSetting(H, W, 1.0, batch, views, txp, rxp, args.grad_mode)

    def step():
        m, s_, q, o, w = params
        ore, oim, radii = cr.rasterize_scenes(m, w, o, s_, q, settings)
        loss_fn(ore, target).backward()
        for p_ in params:
            p_.grad = None

    # one receiver per call, scene by scene; loss normalised per spectrum
    pl_params, pl_settings, pl_targets = [], [], []
    r0 = 0
    for sc, d in zip(scs, per):
        prm = [d[k].clone().requires_grad_(True) for k in ("means", "scales", "rots", "opac", "mlp")]
        for r in range(sc.R):
            pl_params.append(prm)
            pl_settings.append(cr.GaussianRasterizerSetting(H, W, 1.0, d["views"][r], torch.eye(4, device=dev),
                                                            d["tx"], d["rx"][r], args.grad_mode))
            pl_targets.append(target[r0 + r:r0 + r + 1])
        r0 += sc.R

    def step_plain():
        plain_step(pl_params, pl_settings, pl_targets, batch.R)
        for prm in pl_params:
            for p_ in prm:
                p_.grad = None

    pairs = [(s, r) for s, sc in enumerate(scs) for r in range(sc.R)]

    def single(i):
        s, r = pairs[(i * 7919) % len(pairs)]
        d = per[s]
        wr.rasterize_gaussians(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], H, W, 1.0, d["views"][r],
                               None, d["tx"], d["rx"][r])

    def count():
        m, s_, q, o, w = packed
        return wr.forward_scenes(m, w, o, s_, q, H, W, 1.0, views, txp, rxp, batch)[0]

    def tune():
        m, s_, q, o, w = packed
        with torch.no_grad():
            out = wr.forward_scenes(m, w, o, s_, q, H, W, 1.0, views, txp, rxp, batch)
        return wr.autotune(out, grad_mode=args.grad_mode, scenes=batch)

    def graphed():
        """The same scene step captured in a HIP graph; tile lists sized for
        1.25 K."""
        import gsparc_train as T
        cap = int(count() * 1.25) + 1024
        live = [x.clone() for x in packed]
        g = T.GraphedSceneStep(*live, views, txp, rxp, target, H, W, batch, capacity=cap,
                               grad_mode=args.grad_mode)
        return g, live

    def cpu():
        # the smallest scene's receivers, in order, up to the sample size
        order = np.argsort([sc.P for sc in scs])
        sample, tg = [], []
        offs = np.cumsum([0] + [sc.R for sc in scs])
        tcpu = target.cpu().numpy()
        for s in order:
            for r in range(scs[s].R):
                if len(sample) < args.cpu_receivers:
                    sample.append((scs[s], r))
                    tg.append(tcpu[offs[s] + r])
        return cpu_oracle_sample(args, sample, tg, f"{len(sample)} spectra of the smallest scenes "
                                                   f"(P = {sorted(sc.P for sc in scs)[:3]}...)")

    return Workload(
        step=step, step_plain=step_plain, single=single, count=count, cpu=cpu, graphed=graphed, tune=tune,
        spectra=batch.R,
        Ptot=batch.Ptot,
        Nrec=batch.Nrec, H=H, W=W, traffic_key=None,
        data="synthetic (seeded: 16 rooms, 70% wall-clustered Gaussians, receivers uniform in each room)",
        config={"workload": "uneven scene batch, fwd+bwd, 0.8*L1 + 0.2*(1-SSIM)", "scenes": len(scs),
                "gaussians": [sc.P for sc in scs], "receivers": [sc.R for sc in scs],
                "total_gaussians": batch.Ptot, "total_receivers": batch.R, "map": f"{H}x{W}",
                "emission": "per-Gaussian MLP 3->32->2", "grad_mode": args.grad_mode})


def setup_csi(args, dev):
    """50k Gaussians (config2 generator), 32 receivers, 52-subcarrier maps."""
    import numpy as np
    import torch
    import cuda_rasterizer as cr
    import gsparc_scenes as S
    import gsparc_train as T
    import wireless_rasterizer as wr
    P, R = args.gaussians, args.csi_receivers
    sc = S.config2(P=P, R=R)
    tgt_sc = S.config2(seed=3, P=P, R=R)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    means, scales, rots, opac, mlp = (t(sc.means), t(sc.scales), t(sc.rotations), t(sc.opacities), t(sc.mlp))
    views, txp, rxp = t(sc.views), t(sc.tx_pos), t(sc.rx_pos)
    with torch.no_grad():
        target = wr.forward_csi(t(tgt_sc.means), t(tgt_sc.mlp), t(tgt_sc.opacities), t(tgt_sc.scales),
                                t(tgt_sc.rotations), sc.H, sc.W, 1.0, views, txp, rxp)[1].clone()
    params = [x.clone().requires_grad_(True) for x in (means, scales, rots, opac, mlp)]
    settings = cr.CsiSetting(sc.H, sc.W, 1.0, views, txp, rxp)

    def step():
        m, s_, q, o, w = params
        out, radii = cr.rasterize_csi(m, w, o, s_, q, settings)
        T.CSILoss.apply(out, target).backward()
        for p_ in params:
            p_.grad = None

    plain_settings = [cr.CsiSetting(sc.H, sc.W, 1.0, views[r:r + 1], txp, rxp[r:r + 1]) for r in range(R)]

    def step_plain():
        m, s_, q, o, w = params
        for r in range(R):
            out, radii = cr.rasterize_csi(m, w, o, s_, q, plain_settings[r])
            (T.CSILoss.apply(out, target[r:r + 1]) / R).backward()
        for p_ in params:
            p_.grad = None

    def single(i):
        r = i % R
        wr.forward_csi(means, mlp, opac, scales, rots, sc.H, sc.W, 1.0, views[r:r + 1], txp, rxp[r:r + 1])

    def count():
        return wr.forward_csi(means, mlp, opac, scales, rots, sc.H, sc.W, 1.0, views, txp, rxp)[0]

    def pair_count_step():
        """The single-channel blend forward and backward on the same geometry:
        they walk exactly the pairs the multi-frequency kernels walk (alpha, T
        and termination do not depend on the subcarrier), and their counting
        variants count them."""
        with torch.no_grad():
            out = wr.forward_batched(means, mlp, opac, scales, rots, sc.H, sc.W, 1.0, views, txp, rxp)
            K_, ore, oim, radii, geom, binbuf, img = out
            g = torch.ones_like(ore)
            wr.backward_batched(means, mlp, radii, opac, scales, rots, 1.0, views, g, g, txp, rxp, geom, K_, binbuf,
                                img, grad_mode="exact")

    def graphed(deterministic=False):
        """The same step (CSI forward, CSI loss, exact adjoint) captured in a
        HIP graph; tile lists sized for 1.25 K."""
        cap = int(count() * 1.25) + 1024
        live = [x.clone() for x in (means, scales, rots, opac, mlp)]
        g = T.GraphedCsiStep(*live, views, txp, rxp, target.contiguous(), sc.H, sc.W, capacity=cap)
        return g, live

    def cpu():
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        nr = min(args.cpu_receivers, R, 8)
        tg = target[:nr].cpu().numpy()
        return cpu_baseline(args, lambda n: cpu_run(n, tg), nr, min(2, nr),
                            f"{nr} of the {R} receivers, 52-subcarrier maps")

    def cpu_run(nr, tg):
        import oracle as O
        acc = O.Grads(P)
        for r in range(nr):
            pre, plen, vals, ranges, fwd = O.render_csi(sc, r)
            x = fwd["out"][None]
            y = tg[r][None]
            ma = np.sqrt((x.astype(np.float32) ** 2).sum(-1)).transpose(0, 3, 1, 2).copy()
            mb = np.sqrt((y ** 2).sum(-1)).transpose(0, 3, 1, 2).copy()
            m_, d1, d2, d3 = O.ssim_fwd(ma, mb, dtype=np.float32)
            dmag = O.ssim_bwd(ma, mb, np.full_like(ma, -0.2 / ma.size), d1, d2, d3, dtype=np.float32)
            dmag = dmag.transpose(0, 2, 3, 1)[..., None]
            mag = ma.transpose(0, 2, 3, 1)[..., None]
            dout = (0.8 / x.size) * np.sign(x - y) + np.where(mag > 0, dmag * x / np.where(mag > 0, mag, 1), 0)
            O.backward_csi(sc, pre, plen, vals, ranges, fwd, dout[0].astype(np.float32), r, grads=acc)

    return Workload(
        step=step, step_plain=step_plain, single=single, count=count, cpu=cpu, graphed=graphed, tune=None,
        pair_count_step=pair_count_step,
        spectra=R, Ptot=P, Nrec=P * R, H=sc.H, W=sc.W, traffic_key=None, csi=True,
        data="synthetic (seeded: 70% wall-clustered Gaussians, receivers uniform in a 6x6x3 m room)",
        config={"workload": "multi-frequency CSI, fwd+bwd, 0.8*L1(complex) + 0.2*(1-SSIM(magnitude))",
                "gaussians": P, "receivers": R, "map": f"{sc.H}x{sc.W}x104", "subcarriers": 52,
                "f0_hz": wr.CSI_F0, "df_hz": wr.CSI_DF, "emission": "per-Gaussian MLP 3->32->2 x phase",
                "grad_mode": "exact"})


MLP_FLOPS_FWD = 2 * (128 * 128 + 128 * 5 + 128)    # per row, in the fused forward kernel
MLP_FLOPS_BWD = 3 * 2 * 128 * 128                  # per row: recompute, dW2, dh1 (MFMA)
PEAK_FP32_MFMA_TFLOPS = 157.3


def setup_config1(args, dev):
    """Single receiver, 20k Gaussians, shared MLP emission (MFMA kernels)."""
    import numpy as np
    import torch
    import cuda_rasterizer as cr
    import gsparc_emission as GE
    import gsparc_scenes as S
    import gsparc_train as T
    import wireless_rasterizer as wr
    sc = S.config1()
    tgt = S.config1(seed=2)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    means, scales, rots, opac = t(sc.means), t(sc.scales), t(sc.rotations), t(sc.opacities)
    views, txp, rxp = t(sc.views), t(sc.tx_pos), t(sc.rx_pos)
    mlp = GE.SharedEmissionMLP(generator=torch.Generator().manual_seed(1)).to(dev)
    with torch.no_grad():
        Et = mlp(t(tgt.means), rxp, txp)
        target = wr.forward_batched(t(tgt.means), None, t(tgt.opacities), t(tgt.scales), t(tgt.rotations), sc.H,
                                    sc.W, 1.0, views, txp, rxp, emission=Et)[1].clone()
    params = [x.clone().requires_grad_(True) for x in (means, opac)]
    settings = cr.BatchedSetting(sc.H, sc.W, 1.0, views, txp, rxp, "exact")

    def step():
        m, o = params
        E = mlp(m, rxp, txp)
        ore, oim, radii = cr.rasterize_gaussians_batched(m, None, o, scales, rots, settings, emission=E)
        T.L1SSIMLoss.apply(ore, target).backward()
        for p_ in list(params) + list(mlp.parameters()):
            p_.grad = None

    def single(i):
        with torch.no_grad():
            E = mlp(means, rxp, txp)
            wr.forward_batched(means, None, opac, scales, rots, sc.H, sc.W, 1.0, views, txp, rxp, emission=E)

    def count():
        with torch.no_grad():
            E = mlp(means, rxp, txp)
        return wr.forward_batched(means, None, opac, scales, rots, sc.H, sc.W, 1.0, views, txp, rxp, emission=E)[0]

    def graphed(deterministic=False):
        """The same step (MLP forward, rasterizer forward, fused loss,
        backward, MLP backward) captured in a HIP graph; tile lists sized for
        1.25 K."""
        cap = int(count() * 1.25) + 1024
        live = [x.clone() for x in (means, scales, rots, opac)]
        weights = [p_.detach().clone().contiguous() for p_ in mlp.parameters()]
        g = T.GraphedSharedMlpStep(*live, weights, views, txp, rxp, target, sc.H, sc.W, capacity=cap)
        return g, (live, weights)

    def mlp_rows():
        """The MLP alone at 50k Gaussians x 64 receivers (3.2M rows), fwd+bwd."""
        import gsparc_hip as gh
        s2 = S.config2()
        mu = t(s2.means).requires_grad_(True)
        rx2, tx2 = t(s2.rx_pos), t(s2.tx_pos)
        w = torch.randn((s2.R, s2.P), device=dev)
        for _ in range(2):
            (mlp.amplitude(mu, rx2, tx2) * w).sum().backward()
        torch.cuda.synchronize()
        gh.profile_enable(True, timing=True, counting=False)
        n = 5
        for _ in range(n):
            (mlp.amplitude(mu, rx2, tx2) * w).sum().backward()
        torch.cuda.synchronize()
        kt = gh.profile_collect()
        gh.profile_enable(False)
        rows = s2.R * s2.P
        f_ms = kt["mlp1_fwd"][0] / kt["mlp1_fwd"][1]
        b_ms = kt["mlp1_bwd"][0] / kt["mlp1_bwd"][1]
        ft = MLP_FLOPS_FWD * rows / (f_ms * 1e-3) / 1e12
        bt = MLP_FLOPS_BWD * rows / (b_ms * 1e-3) / 1e12
        for p_ in mlp.parameters():
            p_.grad = None
        return {"rows": rows, "fwd_ms": round(f_ms, 4), "bwd_ms": round(b_ms, 4), "fwd_tflops": round(ft, 2),
                "bwd_tflops": round(bt, 2), "peak_tflops": PEAK_FP32_MFMA_TFLOPS,
                "fwd_frac": round(ft / PEAK_FP32_MFMA_TFLOPS, 4), "bwd_frac": round(bt / PEAK_FP32_MFMA_TFLOPS, 4),
                "flops_per_row": {"fwd": MLP_FLOPS_FWD, "bwd": MLP_FLOPS_BWD}}

    return Workload(
        step=step, step_plain=None, single=single, count=count, cpu=lambda: None, graphed=graphed, tune=None,
        spectra=1, Ptot=sc.P, Nrec=sc.P, H=sc.H, W=sc.W, traffic_key=None, mlp_rows=mlp_rows,
        data="synthetic (seeded: 20k Gaussians uniform in a 6x6x3 m room, receiver at the centre)",
        config={"workload": "single-receiver training step, shared 68->128->128->1 emission MLP (FP32 MFMA), "
                            "0.8*L1 + 0.2*(1-SSIM)", "gaussians": sc.P, "receivers": 1, "map": f"{sc.H}x{sc.W}",
                "grad_mode": "exact"})


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the receiver-sharded multi-GPU "
              "path is a design note, not part of this build.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import gsparc_hip as gh

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    gh.lib()
    gh.set_option("render_fwd_ppl", args.fwd_ppl)
    gh.set_option("render_bwd_ppl", args.bwd_ppl)
    gh.set_option("render_order", args.order)
    gh.set_option("render_fwd_depth", args.fwd_depth)
    gh.set_option("render_bwd_depth", args.bwd_depth)
    wl = {"receivers": setup_receivers, "scenes": setup_scenes, "csi": setup_csi,
          "config1": setup_config1}[args.workload](args, dev)
    R, Ptot, Nrec = wl.spectra, wl.Ptot, wl.Nrec
    tuned = None
    if wl.tune is not None and not args.no_tune and not (args.fwd_ppl or args.bwd_ppl or args.order or
                                                             args.fwd_depth or args.bwd_depth):
        tuned = wl.tune()
        log(args, f"autotuned blend mapping: {tuned}")

    def timed(fn, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    log(args, f"warmup {args.warmup} batched steps ({args.workload}: {R} spectra, {Ptot} Gaussians)")
    for _ in range(args.warmup):
        wl.step()
    torch.cuda.synchronize()
    log(args, f"timing {args.steps} eager batched steps")
    el_eager = timed(wl.step, args.steps)
    eager_value = R * args.steps / el_eager
    log(args, f"eager batched: {1e3 * el_eager / args.steps:.2f} ms/step, {eager_value:.1f} spectra/s")
    graph_value = None
    gstep = None
    if wl.graphed is not None and not args.eager:
        gstep, _live = wl.graphed()
        for _ in range(args.warmup):
            gstep.replay()
        torch.cuda.synchronize()
        if gstep.overflowed():
            raise RuntimeError("graphed step: tile lists exceeded their capacity")
        log(args, f"timing {args.steps} graph-captured batched steps")
        el_graph = timed(gstep.replay, args.steps)
        graph_value = R * args.steps / el_graph
        log(args, f"graphed batched: {1e3 * el_graph / args.steps:.2f} ms/step, {graph_value:.1f} spectra/s")
    el = el_graph if graph_value is not None else el_eager
    ms_per_step = 1e3 * el / args.steps
    value = R * args.steps / el

    step_lat = []
    if args.workload == "scenes":
        # every spectrum of a batch is ready when its step ends
        for _ in range(max(5, args.steps // 2)):
            step_lat.append(1e3 * timed(wl.step, 1))

    plain_value = None
    if args.plain_steps > 0 and wl.step_plain is not None:
        wl.step_plain()
        log(args, f"timing {args.plain_steps} plain steps (one receiver per call)")
        el_plain = timed(wl.step_plain, args.plain_steps)
        plain_value = R * args.plain_steps / el_plain
        log(args, f"plain: {1e3 * el_plain / args.plain_steps:.2f} ms/step, {plain_value:.1f} spectra/s")

    # single-spectrum forward latency (one receiver, reference API)
    lat = []
    with torch.no_grad():
        for i in range(args.latency_iters + 5):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            wl.single(i)
            torch.cuda.synchronize()
            if i >= 5:
                lat.append(1e3 * (time.perf_counter() - t0))
    lat = np.asarray(lat if lat else [float("nan")])
    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(args, f"calibration copy of {args.copy_mib} MiB done")

    # per-kernel times with HIP events on the launch stream around every
    # kernel of the timed step: the graph's own launch sequence run without
    # the graph (the same kernels with the same arguments), else the eager
    # step; then (separately: the counting variants are slower) the blend
    # kernels' pair counts.  After the host-bound plain and latency runs the
    # clocks need about 40 steps to come back to where the timed steps ran
    # (with 5, the blend kernels read 7-10% slower and the kernels' sum
    # exceeded the graphed step by 7%).
    prof_step = gstep.run_eager if gstep is not None else wl.step
    for _ in range(max(40, args.warmup)):
        prof_step()
    torch.cuda.synchronize()
    nprof = 5
    gh.profile_enable(True, timing=True, counting=False)
    for _ in range(nprof):
        prof_step()
    torch.cuda.synchronize()
    kt = gh.profile_collect()
    gh.profile_enable(True, timing=False, counting=True)
    count_step = getattr(wl, "pair_count_step", None) or prof_step
    for _ in range(nprof):
        count_step()
    torch.cuda.synchronize()
    cnt = {k: v / nprof for k, v in gh.profile_counts().items()}
    gh.profile_enable(False)
    if gstep is not None:
        gstep.close()
        del gstep, _live
    with torch.no_grad():
        K = wl.count()
    N = wl.W * wl.H
    T = ((wl.W + 15) // 16) * ((wl.H + 15) // 16)
    # algorithmic HBM bytes per launch (each datum read or written once)
    work = {
        # 44 B per tile-list instance (index 4, splat record 32, emission 8)
        # + 16 B per pixel (out re/im, final T, contributor count)
        "render_fwd": ("hbm", 44.0 * K + 16.0 * N * R, "GB/s"),
        # + 16 B per pixel read (final T, count, dL/dout re/im) + 32 B per
        # (receiver, Gaussian) accumulated gradient record
        "render_bwd": ("hbm", 44.0 * K + 16.0 * N * R + 32.0 * Nrec, "GB/s"),
        # + 56 B per (Gaussian, receiver): splat record 32, emission 8,
        # radius 4, depth 4, path length 4, packed tile rect 4
        "preprocess_fwd": ("hbm", Ptot * (820 + 24) + 56 * Nrec, "GB/s"),
        # + 84 B per (Gaussian, receiver): gradient record 32, conic 12,
        # radius 4, dL/dmean2D out 12, MLP outputs 8, dL/dz 8, partials 8
        "preprocess_bwd": ("hbm", Ptot * 1690 + 84 * Nrec, "GB/s"),
        # depth and radius read twice (the digit counts, then the first
        # pass), 3 passes reading and all 4 writing (key, index) pairs, and
        # the packed rects carried into depth order by the last pass
        "bin_depth_sort": ("hbm", 80.0 * Nrec, "GB/s"),
        # depth-ordered rects in, one count per (256-Gaussian chunk, tile) out
        "bin_chunk_count": ("hbm", 4.0 * Nrec + 4.0 * T * (Nrec / 256.0), "GB/s"),
        # sorted index and rect per (Gaussian, receiver) in, one index per instance out
        "bin_scatter": ("hbm", 4.0 * K + 8.0 * Nrec, "GB/s"),
        "ssim_fwd": ("hbm", 8 * N * R + 16 * N * R, "GB/s"),
        "ssim_bwd": ("hbm", 28 * N * R, "GB/s"),
    }
    if getattr(wl, "mlp_rows", None):
        work["mlp1_fwd"] = ("mfma", MLP_FLOPS_FWD * Nrec, "TFLOP/s")
        work["mlp1_bwd"] = ("mfma", MLP_FLOPS_BWD * Nrec, "TFLOP/s")
    if getattr(wl, "csi", False):
        # 104 floats per pixel out (forward) and dL/dout in (backward); the
        # SSIM runs on 52 magnitude planes per receiver
        NC = 52
        work.update({
            "render_fwd_csi": ("hbm", 44.0 * K + (8 * NC + 8) * N * R, "GB/s"),
            "render_bwd_csi": ("hbm", 44.0 * K + (8 * NC + 8) * N * R + 36.0 * Nrec, "GB/s"),
            "csi_mag": ("hbm", 24.0 * NC * N * R, "GB/s"),
            "csi_grad": ("hbm", 32.0 * NC * N * R, "GB/s"),
            "ssim_fwd": ("hbm", 24.0 * NC * N * R, "GB/s"),
            "ssim_bwd": ("hbm", 28.0 * NC * N * R, "GB/s"),
        })
    # The blend kernels are bound by the vector ALU: their algorithmic work is
    # the reference's FP32 operations over the pairs its loops visit (counted
    # by the kernels' counting variants), against the published FP32 vector
    # peak; their HBM rate is kept beside it.
    valu = {"render_fwd": (cnt["fwd_pairs"] * FWD_FLOP_PER_PAIR, cnt["fwd_pairs"], cnt["fwd_contrib"]),
            "render_bwd": (cnt["bwd_pairs"] * BWD_FLOP_PER_PAIR + cnt["bwd_contrib"] * BWD_FLOP_PER_CONTRIB,
                           cnt["bwd_pairs"], cnt["bwd_contrib"])}
    if getattr(wl, "csi", False):
        # the same walk with 52 complex channels: the forward accumulates
        # C_k += w e_k (104 FMAs per contributing pair), the adjoint forms
        # Z = sum_k conj(G_k) u_k (4 FMAs per subcarrier) and Y = sum_k f_k p_k
        # (2 FMAs per subcarrier) per contributing pair
        valu = {"render_fwd_csi": (cnt["fwd_pairs"] * FWD_FLOP_PER_PAIR + cnt["fwd_contrib"] * CSI_FWD_FLOP_PER_CONTRIB,
                                   cnt["fwd_pairs"], cnt["fwd_contrib"]),
                "render_bwd_csi": (cnt["bwd_pairs"] * BWD_FLOP_PER_PAIR + cnt["bwd_contrib"] * CSI_BWD_FLOP_PER_CONTRIB,
                                   cnt["bwd_pairs"], cnt["bwd_contrib"])}
    traffic = measured_traffic(args, *wl.traffic_key) if wl.traffic_key else {}
    kernels = {}
    for name, (ms, n) in kt.items():
        avg = ms / max(n, 1)
        entry = {"ms_per_launch": round(avg, 4), "launches_per_step": n / nprof}
        if name in work:
            bound, amount, unit = work[name]
            if unit == "TFLOP/s":
                ach = amount / (avg * 1e-3) / 1e12
                frac = ach / PEAK_FP32_MFMA_TFLOPS
            else:
                ach = amount / (avg * 1e-3) / 1e9
                frac = ach / PEAK_HBM_GBS
            entry.update(bound=bound, achieved=round(ach, 3 if unit == "TFLOP/s" else 1), unit=unit,
                         frac=round(frac, 4))
        if name in valu:
            fl, pairs, contrib = valu[name]
            gbs, gfrac = entry.get("achieved"), entry.get("frac")
            ach = fl / (avg * 1e-3) / 1e12
            entry.update(bound="valu", achieved=round(ach, 3), unit="TFLOP/s", frac=round(ach / PEAK_FP32_TFLOPS, 4),
                         flop_per_launch=fl, pairs_per_launch=pairs, contributing_pairs_per_launch=contrib,
                         hbm_gbs=gbs, hbm_frac=gfrac)
        if name in traffic:
            entry["traffic"] = traffic[name]
        kernels[name] = entry
    dom = max(kt.items(), key=lambda kv: kv[1][0])[0]
    d = kernels.get(dom, {})
    peak = {"TFLOP/s": PEAK_FP32_TFLOPS if d.get("bound") == "valu" else PEAK_FP32_MFMA_TFLOPS}.get(d.get("unit"),
                                                                                                 PEAK_HBM_GBS)
    roof = {"kernel": dom, "bound": d.get("bound"), "achieved": d.get("achieved"), "peak": peak,
            "unit": d.get("unit"), "frac": d.get("frac"), "traffic": traffic.get(dom),
            "ms_per_launch": d.get("ms_per_launch")}
    if dom in valu:
        roof.update(flop_per_launch=d["flop_per_launch"], pairs_per_launch=d["pairs_per_launch"],
                    contributing_pairs_per_launch=d["contributing_pairs_per_launch"],
                    flop_per_pair=FLOP_PER_PAIR_TEXT[dom],
                    hbm_gbs=d.get("hbm_gbs"), hbm_frac=d.get("hbm_frac"))
    if roof["traffic"] is not None:
        roof["traffic_source"] = TRAFFIC_FILE

    train_iter = None
    if getattr(wl, "trainer", None) is not None and not args.eager:
        tr = wl.trainer()
        for _ in range(args.warmup):
            tr.step()
        torch.cuda.synchronize()
        log(args, f"timing {args.steps} graph-captured training iterations")
        el_tr = timed(tr.step, args.steps)
        if tr.overflowed() or tr.nan_centre():
            raise RuntimeError("training iterations: tile-list overflow or NaN centres")
        tr.close()
        # the surround's own kernels, timed on an eager trainer
        tr2 = wl.trainer(use_graphs=False)
        for _ in range(3):
            tr2.step()
        torch.cuda.synchronize()
        gh.profile_enable(True, timing=True, counting=False)
        for _ in range(nprof):
            tr2.step()
        torch.cuda.synchronize()
        kt2 = gh.profile_collect()
        gh.profile_enable(False)
        tr2.close()
        train_iter = {"ms_per_iteration": round(1e3 * el_tr / args.steps, 3),
                      "spectra_per_s": round(R * args.steps / el_tr, 2),
                      "step": "hip graph: forward, confidence-weighted loss, backward, densification statistics, "
                              "Adam with the activations, centre clamp",
                      "surround_ms_per_iteration": {k: round(kt2[k][0] / max(kt2[k][1], 1) *
                                                             (kt2[k][1] / nprof), 4)
                                                    for k in ("adam", "densify_stats", "conf_loss") if k in kt2}}
        log(args, f"training iteration: {train_iter}")

    det_step = None
    if args.workload == "receivers" and not args.eager:
        # the same graphed step with deterministic gradients (fixed-order sums)
        dstep, _dlive = wl.graphed(deterministic=True)
        for _ in range(args.warmup):
            dstep.replay()
        torch.cuda.synchronize()
        el_det = timed(dstep.replay, args.steps)
        if dstep.overflowed():
            raise RuntimeError("deterministic step: tile lists exceeded their capacity")
        for _ in range(3):
            dstep.run_eager()
        torch.cuda.synchronize()
        gh.profile_enable(True, timing=True, counting=False)
        for _ in range(nprof):
            dstep.run_eager()
        torch.cuda.synchronize()
        kd = gh.profile_collect()
        gh.profile_enable(False)
        det_step = {"ms_per_step": round(1e3 * el_det / args.steps, 3),
                    "spectra_per_s": round(R * args.steps / el_det, 2),
                    "step": "hip graph: the timed step with fixed-order gradient sums (grad_mode | "
                            "GSPARC_GRAD_DETERMINISTIC): one wave per tile storing each instance's sums, "
                            "then a per-record gather",
                    "kernels_ms_per_step": {k: round(kd[k][0] / nprof, 4)
                                            for k in ("render_bwd", "det_clear", "det_offsets", "det_gather") if k in kd}}
        log(args, f"deterministic step: {det_step}")
        del dstep, _dlive

    cpu = wl.cpu() if args.cpu_receivers > 0 else None

    metric = {"receivers": METRIC,
              "scenes": "fwd+bwd spectra/sec, 16 uneven scenes (5k-200k Gaussians, 4-64 rx each), 90x360, one batch",
              "csi": "fwd+bwd CSI maps/sec, 50k Gaussians, 90x360x52 complex subcarriers, 32 rx batched",
              "config1": "training steps/sec, 20k Gaussians, 1 rx, shared emission MLP on MFMA"}
    out = {
        "metric": metric[args.workload],
        "value": round(value, 2),
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": wl.data,
        "config": wl.config,
        "step": "hip graph: forward, fused loss, backward" if graph_value is not None else
                "eager: torch autograd over the rasterizer and fused SSIM",
        "eager_spectra_per_s": round(eager_value, 2),
        "plain_spectra_per_s": None if plain_value is None else round(plain_value, 2),
        "batched_over_plain": None if plain_value is None else round(value / plain_value, 2),
        "single_spectrum_fwd_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                   "p99": round(float(np.percentile(lat, 99)), 3)},
        "K_instances": int(K),
        "autotune": tuned,
        "roofline": roof,
        "kernels": kernels,
        "kernels_source": ("HIP events around each kernel of the graph's launch sequence, run without the graph"
                           if graph_value is not None else "HIP events around each kernel of the eager step"),
        "cpu_baseline": cpu,
        "training_iteration": train_iter,
        "deterministic_step": det_step,
    }
    if getattr(wl, "mlp_rows", None):
        out["mlp_mfma"] = wl.mlp_rows()
    if step_lat:
        out["batch_step_ms"] = {"p50": round(float(np.percentile(step_lat, 50)), 3),
                                "p99": round(float(np.percentile(step_lat, 99)), 3)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
