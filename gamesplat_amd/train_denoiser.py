"""Train the learned denoiser on pairs this renderer makes.

python -m gamesplat_amd.train_denoiser [--scenes 40] [--steps-per-scene 60] [--batch 8] [--lr 3e-4]
                                       [--out gamesplat_amd/weights/unet_denoiser.f32]

For each seeded synthetic scene (100k-6M Gaussians, 16-256 clusters,
opacity logit mean 0 or 1) and a few random cameras (radius 10-20, mixed
resolutions), the renderer produces the sorted-alpha image (target) and a
GPS running mean of 1, 2 or 4 samples plus GPS depth (input); random
256x256 crops go to the denoiser's pool, and Adam steps on random batches
of the pool follow.  Most views take a work budget that holds every work
item the frame requests (128 Mi items); --drop-fraction of them take the
16 Mi budget, under which large scenes drop work.  Progress lines report the mean loss; the final
parameters (float32, gs_denoiser parameter layout) are written to --out,
with a JSON summary next to it.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import gamesplat_amd as gs  # noqa: E402
from gamesplat_amd import abi, scenes  # noqa: E402

DEFAULT_OUT = Path(__file__).resolve().parent / "weights" / "unet_denoiser.f32"
SIZES = [(640, 480), (960, 540), (1280, 720), (1920, 1080)]
COUNTS = [100_000, 250_000, 500_000, 1_000_000, 2_000_000, 4_000_000, 6_000_000]
FULL_BUDGET = 128 * 1024 * 1024
DROP_BUDGET = 16 * 1024 * 1024


def random_camera(rng, size):
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    pose = scenes.look_at_origin_pose((d * rng.uniform(10.0, 20.0)).astype(np.float32))
    return scenes.camera_from_pose(pose, scenes.FOCAL_OVER_HEIGHT, size[0], size[1])


def render_pair(ctx, cam, spp, frame_seed):
    """Sorted alpha, then an spp-sample GPS running mean of the same camera."""
    ctx.render_alpha(cam)
    for f in range(1, spp + 1):
        ctx.gps_sample(frame_seed=frame_seed + f, supersample=1)
        ctx.accumulate(f)


def psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0.0 else 10.0 * np.log10(1.0 / mse)


def evaluate(ctx, den, g, seed, views, spp, size):
    """Mean PSNR against sorted alpha of the GPS mean and of its denoised
    image over `views` random cameras of scene g."""
    rng = np.random.default_rng(seed)
    ctx.upload(g)
    pa, pd = [], []
    for k in range(views):
        cam = random_camera(rng, size)
        render_pair(ctx, cam, spp, 900000 + 16 * k)
        alpha = ctx.read_image(abi.GS_IMAGE_ALPHA_RGBA16F)
        acc = ctx.read_image(abi.GS_IMAGE_ACCUM_RGBA32F)
        den.denoise()
        out = den.read_image(size[0], size[1])
        pa.append(psnr(acc, alpha))
        pd.append(psnr(out, alpha))
    return float(np.mean(pa)), float(np.mean(pd))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=40)
    ap.add_argument("--views", type=int, default=4)
    ap.add_argument("--crops", type=int, default=8)
    ap.add_argument("--steps-per-scene", type=int, default=60)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--lr-final", type=float, default=None,
                    help="learning rate of the last scene (linear from --lr); default: --lr throughout")
    ap.add_argument("--seed", type=int, default=0x5EED7000)
    ap.add_argument("--init", type=str, default="", help="start from these parameters")
    ap.add_argument("--out", type=str, default=str(DEFAULT_OUT))
    ap.add_argument("--budget-s", type=float, default=900.0, help="stop adding scenes after this many seconds")
    ap.add_argument("--drop-fraction", type=float, default=0.15,
                    help="share of views rendered with the 16 Mi work budget")
    args = ap.parse_args(argv)

    rng = np.random.default_rng(args.seed)
    t0 = time.time()
    log = []
    with gs.Context(0) as ctx, gs.Denoiser(ctx, seed=args.seed) as den:
        if args.init:
            den.set_params(np.fromfile(args.init, dtype=np.float32))
        steps = 0
        for s in range(args.scenes):
            if time.time() - t0 > args.budget_s:
                break
            count = int(rng.choice(COUNTS))
            clusters = int(rng.choice([16, 64, 256]))
            opacity_mean = float(rng.choice([0.0, 1.0]))
            g = gs.generate_scene(args.seed + s, count, clusters, opacity_mean)
            ctx.upload(g)
            for v in range(args.views):
                size = SIZES[int(rng.integers(len(SIZES)))]
                spp = int(rng.choice([1, 1, 2, 4]))
                ctx.set_gps_budget(DROP_BUDGET if rng.uniform() < args.drop_fraction else FULL_BUDGET)
                render_pair(ctx, random_camera(rng, size), spp, (s * 64 + v) * 16)
                den.capture(args.crops, args.seed * 31 + s * 64 + v)
            lr = args.lr
            if args.lr_final is not None and args.scenes > 1:
                lr = args.lr + (args.lr_final - args.lr) * s / (args.scenes - 1)
            losses = den.train(args.steps_per_scene, args.batch, lr, args.seed + 7 * s)
            steps += args.steps_per_scene
            rec = {"scene": s, "lr": lr, "gaussians": count, "clusters": clusters, "opacity_mean": opacity_mean,
                   "steps": steps, "loss": float(losses.mean()), "pool": den.pool_size(),
                   "elapsed_s": round(time.time() - t0, 1)}
            log.append(rec)
            print(json.dumps(rec), flush=True)
        params = den.get_params()
        # held-out scenes (seeds outside the training range)
        evals = []
        ctx.set_gps_budget(FULL_BUDGET)
        for k, (count, clusters, om, size) in enumerate([(1_000_000, 64, 0.0, (1280, 720)),
                                                         (500_000, 256, 1.0, (1280, 720)),
                                                         (6_000_000, 256, 1.0, (1920, 1080))]):
            g = gs.generate_scene(args.seed + 100000 + k, count, clusters, om)
            for spp in (1, 4):
                pa, pd = evaluate(ctx, den, g, args.seed + 200000 + k, 2, spp, size)
                evals.append({"gaussians": count, "size": list(size), "spp": spp, "budget": FULL_BUDGET,
                              "psnr_gps_mean": round(pa, 2), "psnr_denoised": round(pd, 2)})
                print(json.dumps(evals[-1]), flush=True)
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    params.astype(np.float32).tofile(out)
    summary = {"param_count": int(params.size), "steps": steps, "batch": args.batch, "lr": args.lr,
               "lr_final": args.lr_final,
               "gaussians": [COUNTS[0], COUNTS[-1]], "drop_fraction": args.drop_fraction,
               "scenes": len(log), "final_loss": log[-1]["loss"] if log else None, "held_out": evals,
               "seconds": round(time.time() - t0, 1)}
    out.with_suffix(".json").write_text(json.dumps(summary, indent=1) + "\n")
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
