"""The benchmark: TGV primal-dual iterations on one MI355X, as Mpixel-iterations per second.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c4] [--precision single|mixed|double]

A step is one iteration of the method over the whole canvas of the workload (all its
channels). The default workload is the 2048 x 2048 RGB picture of seed 0, JFIF YCbCr
4:2:0 at quality 75 (IJG tables), TGV with the default alpha1 = 1, alpha0 = 2, in
binary32. W warm-up iterations run in a solve of their own; then one solve of exactly K
iterations is timed, its records only at its end (--record-every 0 --tolerance 0),
bracketed by torch.cuda.synchronize(). value counts the samples of every channel times K
over the iterations' own time (HIP events on the solve's stream around each iteration,
records and setup excluded, as the reference's throughput does); ms_per_step is that time
per iteration, and the wall time of the whole timed solve is printed to stderr.

roofline: the iteration's kernels read and write HBM; the algorithmic bytes of a sample
and iteration are 2 x the bytes of the state's fields (x, w, p, r) + 2 x the coefficients
per sample (each int16 level read once), and achieved = those bytes of one iteration over
its measured duration, against MI355X's 8 TB/s. cpu_baseline: the CPU oracle (binary64,
one thread) on a bounded sample of the same kind of workload, timed on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for path in (ROOT, ROOT / "jpeg-unround_amd"):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))

METRIC = "TGV Mpixel-iters/s + %HBM roofline; batched vs plain images/s at equal PSNR"
HBM_PEAK_GBS = 8000.0


def log(*parts: object) -> None:
    print(*parts, file=sys.stderr, flush=True)


def workload(name: str):
    from unround import synthetic

    if name == "c2":
        picture = synthetic.colour_picture(2048, 2048, 0)
        return synthetic.encode_colour420(picture, 75), "tgv", "C2: 2048x2048 RGB YCbCr 4:2:0 q75, TGV"
    if name == "c1":
        picture = synthetic.grey_picture(256, 256, 0)
        return synthetic.encode_grey(picture, 50), "tv", "C1: 256x256 grey q50, TV"
    if name == "c4":
        picture = synthetic.grey_picture(4096, 4096, 0)
        return synthetic.encode_grey(picture, 40), "tgv", "C4: 4096x4096 grey q40, TGV"
    raise SystemExit(f"unknown config {name}")


def field_bytes(method: str, precision: str) -> float:
    """Bytes of one sample's state: x (and w) at the iterate's size, p (and r) at the dual storage's."""
    iterate = 8 if precision == "double" else 4
    dual = 2 if precision == "mixed" else iterate
    return iterate + 2 * dual if method == "tv" else 3 * iterate + 5 * dual


def cpu_baseline(method: str, budget_seconds: float = 12.0) -> dict:
    """The oracle on a 512 x 512 sample of the workload's kind, about budget_seconds of work."""
    import numpy as np

    from oracle import oracle
    from unround import synthetic

    if method == "tv":
        picture = synthetic.grey_picture(512, 512, 0)
        image = synthetic.encode_grey(picture, 50)
        kind = "512x512 grey q50 TV"
    else:
        picture = synthetic.colour_picture(512, 512, 0)
        image = synthetic.encode_colour420(picture, 75)
        kind = "512x512 RGB YCbCr 4:2:0 q75 TGV"
    components = [oracle.Component(np.asarray(c.coefficients), np.asarray(c.quant_table).reshape(64),
                                   c.h_samp_factor, c.v_samp_factor) for c in image.components]

    def run(iterations: int) -> tuple[float, int]:
        options = ["--method", method, "--iterations", str(iterations), "--tolerance", "0", "--record-every", "0"]
        began = time.perf_counter()
        result = oracle.solve(image.height, image.width, int(image.color_space), components, options)
        return time.perf_counter() - began, int(np.prod(result.canvas.shape))

    short, _ = run(2)
    longer, _ = run(6)
    per_iteration = max((longer - short) / 4, 1e-6)
    iterations = max(4, int(budget_seconds / per_iteration))
    seconds, samples = run(iterations)
    return {
        "value": round(samples * iterations / seconds / 1e6, 3),
        "unit": "Mpixel-iters/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{kind}, {iterations} iterations in {seconds:.1f} s (records and setup included)",
    }


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=500)
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--config", default="c2")
    parser.add_argument("--precision", default="single", choices=["single", "mixed", "double"])
    parser.add_argument("--no-cpu-baseline", action="store_true")
    args = parser.parse_args()
    if args.gpus != 1:
        raise SystemExit("this benchmark runs on one GPU: the path shards by picture across GPUs with no "
                         "communication, and no figure on several GPUs is measured here (--gpus 1)")

    import numpy as np
    import torch

    from unround import native

    image, method, description = workload(args.config)
    log(f"workload {description}; library {native.library_path()}")
    base = ["--method", method, "--tolerance", "0", "--precision", args.precision]
    if args.warmup > 0:
        native.solve(image, base + ["--iterations", str(args.warmup), "--record-every", "0"])
    os.environ["UNROUND_TIMING"] = "1"
    torch.cuda.synchronize()
    began = time.perf_counter()
    decoded = native.solve(image, base + ["--iterations", str(args.steps), "--record-every", "0"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - began
    os.environ.pop("UNROUND_TIMING", None)
    times = decoded.kernel_times
    assert times is not None and times.iterations == args.steps, "the library took no timings"
    channels, rows, columns = decoded.canvas.shape
    samples = channels * rows * columns
    seconds = times.iterations_seconds
    value = samples * args.steps / seconds / 1e6
    coefficients = sum(int(np.prod(c.shape)) for c in decoded.coefficients)
    per_sample = 2 * field_bytes(method, args.precision) + 2 * coefficients / samples
    achieved = per_sample * samples / (seconds / args.steps) / 1e9
    log(f"wall {wall:.3f} s for the timed solve ({args.steps} iterations + setup + final projection); "
        f"iterations {seconds:.4f} s; primal kernels {times.primal_seconds / max(times.primal_launches, 1) * 1e6:.1f} us "
        f"x {times.primal_launches}, dual {times.dual_seconds / max(times.dual_launches, 1) * 1e6:.1f} us "
        f"x {times.dual_launches}")
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpixel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(seconds / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"single": "f32", "mixed": "f32+bf16", "double": "f64"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": description,
            "height": int(image.height),
            "width": int(image.width),
            "channels": int(channels),
            "samples_per_iteration": int(samples),
            "precision": args.precision,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "one iteration (primal_kernel per component + dual_kernel)",
            "bytes_per_sample_iteration": round(per_sample, 3),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(method),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
