"""The benchmark: TGV primal-dual iterations on one MI355X, as Mpixel-iterations per second.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c4|png|c3|c5]
                [--precision single|mixed|double]

A step is one iteration of the method over the whole canvas of the workload (all its
channels). The default workload is the 2048 x 2048 RGB picture of seed 0, JFIF YCbCr
4:2:0 at quality 75 (IJG tables), TGV with the default alpha1 = 1, alpha0 = 2, in
binary32, K = 500. --config c4 is the 4096 x 4096 grey picture at quality 40 with BF16 p
and r (binary32 x and w, binary64 records and final step), K = 1000, unless --precision
or --steps say otherwise. W warm-up iterations run in a solve of their own (first
launches, the dual step's rows measured); then the timed solve runs W untimed iterations
and exactly K timed ones straight after them (so that the K start at the GPU's running
clocks), its records only at its end (--record-every 0 --tolerance 0), bracketed by
torch.cuda.synchronize(). value counts the samples of every channel times K
over the iterations' own time (one pair of HIP events on the solve's stream around the K
iterations, records and setup excluded, as the reference's throughput does); ms_per_step
is that time per iteration, and the wall time of the whole timed solve is printed to
stderr.

roofline: of the dominant kernel, the dual step (dual_kernel, one launch per iteration
over every channel), timed launch by launch with HIP events in a second solve of
min(K, 100) iterations (after W untimed ones, as in the timed solve). Its algorithmic
bytes per channel-sample are one read of x-, p, w and r and one write of p, w and r
(60 B in binary32 TGV, 40 B with bfloat16 p and r);
achieved = those bytes of one launch over its mean duration, against MI355X's 8 TB/s.
traffic: the HBM bytes of one launch from rocprofv3 counters (FETCH_SIZE + WRITE_SIZE,
scaled by a copy kernel's reading), where profiles/traffic.json holds them for this
workload and the dual step's rows per wavefront that the solve measured best and took
(dual_rows). iteration_achieved is the same for a whole iteration (2 x the state's bytes +
2 x the coefficients per sample, each int16 level read once).

accuracy: the timed solve's result against the same solve in binary64 on the GPU (largest
difference of the canvases, PSNR against the uncompressed picture, requantization).

cpu_baseline: the CPU oracle (binary64; its iterations threaded over bands of rows, the
same bits on any count of threads) on the workload itself, as the reference's throughput
harness times its solvers (rust/unround/tests/throughput.rs:21-61): fixed iterations with
no tolerance and no records between them, the least of 5 single iterations, setup and
records excluded; on 1 thread and on every core this process may use (value), with nproc
and the CPU model.

batch (the default run, and --config c3): the mixed batch of 64 pictures (C3) solved
together (unround_solve_batch) and one at a time (unround_solve), TGV to a relative gap of
1e-4 checked every 10 iterations with the reference's default cap of 10000 TGV iterations
(whatever --steps says): images/s both ways, per-image latency, how many pictures the gap
stopped, whether both ways give the same canvases, and PSNR against binary64 on every
eighth picture. --no-batch leaves it out.

--config png: the same measurement for a picture of rounded samples (unround_solve_samples,
the solve behind a PNG file): 2048 x 2048 RGB levels of 8 bits, TGV in binary32; its
primal kernel is the per-sample proximal step (samples_prox_kernel), the data term 2
bytes a sample.
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
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
    if name == "png":
        return samples_levels(2048, 2048), "tgv", "PNG: 2048x2048 RGB 8-bit rounded samples, TGV"
    raise SystemExit(f"unknown config {name}")


def samples_levels(height: int, width: int):
    """A picture of rounded samples: the synthetic colour picture of seed 0 rounded to 8 bits,
    (height, width, 3) uint16 levels."""
    import numpy as np

    from unround import synthetic

    picture = synthetic.colour_picture(height, width, 0)
    whole = np.floor(picture)
    return np.clip(whole + (picture - whole >= 0.5), 0, 255).astype(np.uint16)


def field_bytes(method: str, precision: str) -> float:
    """Bytes of one sample's state: x (and w) at the iterate's size, p (and r) at the dual storage's."""
    iterate = 8 if precision == "double" else 4
    dual = 2 if precision == "mixed" else iterate
    return iterate + 2 * dual if method == "tv" else 3 * iterate + 5 * dual


def cpu_description() -> dict:
    """nproc, the cores this process may use, and the CPU's model name."""
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    limit = os.environ.get("OMP_NUM_THREADS")
    if limit and limit.isdigit() and int(limit) > 0:
        usable = min(usable, int(limit))
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def cpu_baseline(image, method: str, samples: bool = False, iterations: int = 5) -> dict:
    """The oracle on the workload itself, timed as the reference's throughput harness times its
    solvers: `iterations` iterations with no tolerance and no records between them, the least
    of single iterations (setup and the final record excluded), on 1 thread and on every usable
    core; Mpixel-iterations/s."""
    import numpy as np

    from oracle import oracle

    cpu = cpu_description()
    options = ["--method", method, "--iterations", str(iterations), "--tolerance", "0", "--record-every", "0"]
    if samples:
        levels = np.ascontiguousarray(np.moveaxis(image, -1, 0))
        count = int(levels.size)

        def run() -> float:
            return oracle.solve_samples(levels, 8, options).least_iteration_seconds
    else:
        components = [oracle.Component(np.asarray(c.coefficients), np.asarray(c.quant_table).reshape(64),
                                       c.h_samp_factor, c.v_samp_factor) for c in image.components]
        count = 0

        def run() -> float:
            nonlocal count
            result = oracle.solve(int(image.height), int(image.width), int(image.color_space), components, options)
            count = int(np.prod(result.canvas.shape))
            return result.least_iteration_seconds

    rates = {}
    for threads in (1, cpu["usable"]):
        oracle.set_threads(threads)
        least = run()
        rates[threads] = count * 1.0 / least / 1e6
        log(f"cpu oracle: {threads} thread(s), least iteration {least:.4f} s, {rates[threads]:.2f} Mpixel-it/s")
    oracle.set_threads(0)
    return {
        "value": round(rates[cpu["usable"]], 3),
        "unit": "Mpixel-iters/s",
        "cores": cpu["usable"],
        "kind": "port",
        "single_thread": round(rates[1], 3),
        "nproc": cpu["nproc"],
        "cpu_model": cpu["model"],
        "sample": f"the workload itself ({count} canvas samples), least of {iterations} single iterations on 1 "
                  f"and on {cpu['usable']} threads, binary64, setup and records excluded",
        "reference_published": "TGV colour 1024x1024 4:2:0, Rust, 1 thread: 94 Mpixel-it/s; 4 threads: 183-210 "
                               "(i7-1360P, experiments/results/performance.md:25,33)",
    }


def dual_bytes(method: str, precision: str) -> float:
    """Algorithmic bytes of the dual kernel per channel-sample: it reads x-, p (and w, r) once
    and writes p (and w, r) once."""
    iterate = 8 if precision == "double" else 4
    dual = 2 if precision == "mixed" else iterate
    if method == "tv":
        return iterate + 2 * dual + 2 * dual
    return (iterate + 2 * dual + 2 * iterate + 3 * dual) + (2 * dual + 2 * iterate + 3 * dual)


def measured_traffic(kernel: str, workload_name: str, precision: str, rows: int) -> float | None:
    """HBM bytes per launch of `kernel` from the committed counter summary, if it covers this case
    at the dual step's rows per wavefront that the solve took."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    for entry in json.loads(path.read_text()).get("entries", []):
        if (entry.get("kernel"), entry.get("config"), entry.get("precision"), entry.get("rows")) == (
                kernel, workload_name, precision, rows):
            return float(entry["bytes_per_launch"])
    return None


def batch_workload(count: int = 64):
    """The mixed batch of synthetic.batch_pictures: sides uniform on the multiples of 8 in
    [96, 1920], quality uniform in [30, 90], grey or 4:2:0 colour, each drawn from its own
    seed; with the uncompressed pictures for PSNR."""
    from unround import synthetic

    return synthetic.batch_pictures(count)


def psnr_of(decoded, original) -> float:
    import numpy as np

    from unround import native

    planes = decoded.planes
    ours = planes[0] if planes.shape[0] == 1 else native.to_rgb(planes)
    error = np.mean((np.clip(np.round(ours), 0, 255) - original.astype(np.float64)) ** 2)
    return float(10 * np.log10(255.0 ** 2 / max(error, 1e-30)))


BATCH_CAP = 10000  # the reference's default cap of TGV iterations (cpp/include/unround/pdhg.hpp:45)


def batch_measure(precision: str, cap: int = BATCH_CAP) -> dict:
    """C3: the mixed batch solved together (unround_solve_batch) and one image at a time
    (unround_solve), TGV to a relative gap of 1e-4 checked every 10 iterations, at most `cap`
    iterations; with PSNR against the same pictures solved in binary64 (every eighth)."""
    import numpy as np
    import torch

    from unround import native

    images, originals, kinds = batch_workload()
    options = ["--method", "tgv", "--tolerance", "0", "--relative-tolerance", "1e-4", "--record-every", "10",
               "--iterations", str(cap), "--precision", precision]
    log(f"C3: {len(images)} images, {sum(k == 'colour' for k in kinds)} 4:2:0 colour; options {' '.join(options)}")
    # Warm-up: the library, the kernels and the batch's choice of rows for this shape.
    native.solve_batch(images[:4], options[:-4] + ["--iterations", "20", "--precision", precision])
    native.solve_batch(images, options[:-4] + ["--iterations", "20", "--precision", precision])
    torch.cuda.synchronize()
    began = time.perf_counter()
    together, latencies = native.solve_batch(images, options)
    torch.cuda.synchronize()
    batch_seconds = time.perf_counter() - began
    log(f"C3 batched: {batch_seconds:.3f} s")
    began = time.perf_counter()
    alone = [native.solve(image, options) for image in images]
    torch.cuda.synchronize()
    plain_seconds = time.perf_counter() - began
    log(f"C3 one at a time: {plain_seconds:.3f} s")
    same = all(np.array_equal(a.canvas, b.canvas) for a, b in zip(together, alone))
    picked = list(range(0, len(images), 8))
    psnrs = [psnr_of(d, o) for d, o in zip(together, originals)]
    deltas = [0.0]
    if precision != "double":
        reference, _ = native.solve_batch([images[k] for k in picked], options[:-2] + ["--precision", "double"])
        deltas = [abs(psnrs[k] - psnr_of(r, originals[k])) for k, r in zip(picked, reference)]
    iterations = [d.result.iterations for d in together]
    by_gap = sum(d.result.stop == native.Stop.CONVERGED for d in together)
    channel_samples = [int(np.prod(d.canvas.shape)) for d in together]
    per_sample = 2 * field_bytes("tgv", precision) + 1  # the state each way and the levels
    moved = sum(n * k * per_sample for n, k in zip(channel_samples, iterations))
    log(f"batched {batch_seconds:.3f} s, one at a time {plain_seconds:.3f} s; identical canvases {same}; "
        f"iterations {min(iterations)}..{max(iterations)} (mean {np.mean(iterations):.0f}), {by_gap} stopped by the "
        f"gap; PSNR {min(psnrs):.2f}..{max(psnrs):.2f} dB, largest |dPSNR| vs binary64 {max(deltas):.5f} dB")
    return {
        "workload": "C3: 64 images, sides U{96..1920 step 8}, q U[30,90], grey or 4:2:0, seeds 1..64, TGV to "
                    "relative gap 1e-4 (every 10 iterations)",
        "images": len(images),
        "colour_images": sum(k == "colour" for k in kinds),
        "samples": int(sum(channel_samples)),
        "iteration_cap": cap,
        "precision": precision,
        "images_per_s": round(len(images) / batch_seconds, 3),
        "plain_images_per_s": round(len(images) / plain_seconds, 3),
        "speedup_vs_plain": round(plain_seconds / batch_seconds, 3),
        "batch_seconds": round(batch_seconds, 3),
        "plain_seconds": round(plain_seconds, 3),
        "latency_p50_s": round(float(np.percentile(latencies, 50)), 4),
        "latency_p99_s": round(float(np.percentile(latencies, 99)), 4),
        "stopped_by_gap": int(by_gap),
        "iterations_min_mean_max": [int(min(iterations)), round(float(np.mean(iterations)), 1), int(max(iterations))],
        "identical_to_plain": bool(same),
        "psnr_delta_vs_f64_max_db": round(max(deltas), 6),
        "psnr_reference_images": len(picked),
        "whole_solve_achieved_gbs": round(moved / batch_seconds / 1e9, 1),
        "whole_solve_frac": round(moved / batch_seconds / 1e9 / HBM_PEAK_GBS, 4),
    }


def batch_main(args) -> None:
    """--config c3: the batch's measurement as the line's value (images/s)."""
    batch = batch_measure(args.precision)
    line = {
        "metric": METRIC,
        "value": batch["images_per_s"],
        "unit": "images/s",
        "n_gpus": 1,
        "steps": batch["iteration_cap"],
        "warmup": 1,
        "ms_per_step": round(batch["batch_seconds"] / batch["images"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"single": "f32", "mixed": "f32+bf16", "double": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": batch["workload"], "images": batch["images"], "precision": args.precision},
        "batch": batch,
        "roofline": {
            "bound": "hbm",
            "kernel": "the batch's iterations (prox and dual launches over all images), whole solve",
            "achieved": batch["whole_solve_achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": batch["whole_solve_frac"],
            "traffic": None,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def unrolled_main(args) -> None:
    """C5: 16 grey 512x512 pictures at quality 30, 100 unrolled binary32 TGV iterations, the
    state kept every 10; loss = the mean squared error to the uncompressed pictures; the
    gradients of alpha0, alpha1 and a 512x512 weight map through the library's adjoint."""
    import numpy as np
    import torch

    from unround import synthetic
    from unround.unrolled import UnrolledTGV

    count, size, iterations, every = 16, 512, 100, 10
    originals = [synthetic.grey_picture(size, size, seed) for seed in range(1, count + 1)]
    images = [synthetic.encode_grey(picture, 30) for picture in originals]
    solve = UnrolledTGV(images, ["--method", "tgv"], iterations=iterations, checkpoint_every=every)
    target = torch.tensor(np.stack(originals), dtype=torch.float32, device="cuda")
    weight_map = torch.ones((size, size), dtype=torch.float32, device="cuda", requires_grad=True)
    alpha0 = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    alpha1 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)

    def once():
        canvas = solve(alpha0, alpha1, weight_map)
        loss = ((canvas - target) ** 2).mean()
        torch.cuda.synchronize()
        middle = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        return float(loss.detach()), middle

    once()  # warm-up
    # Five timed forward and backward pairs, each of the two times their median; the
    # gradients reported are the last pair's (one backward each).
    pairs = []
    for _ in range(5):
        alpha0.grad = alpha1.grad = weight_map.grad = None
        torch.cuda.synchronize()
        began = time.perf_counter()
        loss, middle = once()
        ended = time.perf_counter()
        pairs.append((middle - began, ended - middle))
    forward = statistics.median(pair[0] for pair in pairs)
    backward = statistics.median(pair[1] for pair in pairs)
    samples = count * size * size
    # The launches' own durations: a handle made with UNROUND_TIMING=2 (events around every
    # launch on its stream), one forward and backward.
    os.environ["UNROUND_TIMING"] = "2"
    timed = UnrolledTGV(images, ["--method", "tgv"], iterations=iterations, checkpoint_every=every)
    os.environ.pop("UNROUND_TIMING")
    alpha0.grad = alpha1.grad = weight_map.grad = None
    canvas = timed(alpha0, alpha1, weight_map)
    ((canvas - target) ** 2).mean().backward()
    torch.cuda.synchronize()
    times = timed.kernel_times()
    timed.close()
    # Algorithmic bytes a sample: the dual step's adjoint reads w, p, r, x-, tau g grad v, the
    # radius and grad w, grad p, grad r after (68 B) and writes grad x- and grad w, grad p,
    # grad r before (32 B); the prox's adjoint reads grad x-, grad x after and a clip mark and
    # writes grad x and tau g grad v (17 B); the forward dual step 64 B (60 B and the radius),
    # the forward prox 23 B (x, p, the level; x, x-, the mark). The backward: the two adjoints
    # every iteration and the forward again for every run but the last (the forward keeps it).
    back_bytes, prox_back_bytes, dual_bytes_per, prox_bytes = 100, 17, 64, 23
    recomputed = (iterations - 1) // every * every  # the iterations of every run but the last
    backward_bytes = samples * (iterations * (back_bytes + prox_back_bytes) + recomputed * (dual_bytes_per + prox_bytes))
    dual_back_s, dual_back_n = times["dual_back"]
    prox_back_s, prox_back_n = times["prox_back"]
    dual_s, dual_n = times["dual"]
    launch = dual_back_s / dual_back_n
    achieved = back_bytes * samples / launch / 1e9
    log(f"C5 launches: dual step's adjoint {1e6 * launch:.1f} us x {dual_back_n}, prox's adjoint "
        f"{1e6 * prox_back_s / prox_back_n:.1f} us x {prox_back_n}, forward dual step {1e6 * dual_s / dual_n:.1f} us x "
        f"{dual_n}; backward {backward_bytes / samples / iterations:.1f} algorithmic bytes a sample-iteration")
    log(f"C5: loss {loss:.4f}; forward {forward * 1e3:.1f} ms, backward {backward * 1e3:.1f} ms; "
        f"dL/dalpha0 {float(alpha0.grad):.6g}, dL/dalpha1 {float(alpha1.grad):.6g}, "
        f"|dL/dm| {float(weight_map.grad.norm()):.6g}")
    line = {
        "metric": METRIC,
        "value": round(samples * iterations / (forward + backward) / 1e6, 2),
        "unit": "Mpixel-iters/s (forward + backward)",
        "n_gpus": 1,
        "steps": iterations,
        "warmup": 1,
        "ms_per_step": round((forward + backward) / iterations * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "C5: 16 x 512x512 grey q30, 100 unrolled TGV iterations, checkpoint every 10",
                   "pictures": count, "size": size, "iterations": iterations, "checkpoint_every": every},
        "forward_ms": round(forward * 1e3, 2),
        "backward_ms": round(backward * 1e3, 2),
        "loss": round(loss, 6),
        "grad_alpha0": float(alpha0.grad),
        "grad_alpha1": float(alpha1.grad),
        "grad_map_norm": float(weight_map.grad.norm()),
        "roofline": {
            "bound": "hbm",
            "kernel": "unrolled_dual_back_kernel",
            "bytes_per_launch": int(back_bytes * samples),
            "launch_us": round(1e6 * launch, 2),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            # The counted bytes of a launch at the adjoint's 16 rows (none where the
            # environment sets other rows).
            "traffic": None if os.environ.get("UNROUND_BACK_ROWS") else measured_traffic(
                "unrolled_dual_back_kernel", "c5", "single", 16),
            "prox_back_us": round(1e6 * prox_back_s / prox_back_n, 2),
            "prox_back_frac": round(prox_back_bytes * samples / (prox_back_s / prox_back_n) / 1e9 / HBM_PEAK_GBS, 4),
            "backward_bytes_per_sample_iteration": round(backward_bytes / samples / iterations, 1),
            "backward_achieved": round(backward_bytes / backward / 1e9, 1),
            "backward_frac": round(backward_bytes / backward / 1e9 / HBM_PEAK_GBS, 4),
            "backward_over_forward": round(backward / forward, 3),
        },
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=None,
                        help="iterations timed (default: 1000 for c4, as its workload states; 500 otherwise)")
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--config", default="c2")
    parser.add_argument("--precision", default=None, choices=["single", "mixed", "double"],
                        help="default: mixed for c4 (binary32 x and w, BF16 p and r, as its workload states); "
                             "single otherwise")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-batch", action="store_true", help="leave the C3 batch out of the default run")
    args = parser.parse_args()
    if args.precision is None:
        args.precision = "mixed" if args.config == "c4" else "single"
    if args.steps is None:
        args.steps = 1000 if args.config == "c4" else 500
    if args.gpus != 1:
        raise SystemExit("this benchmark runs on one GPU: the path shards by picture across GPUs with no "
                         "communication, and no figure on several GPUs is measured here (--gpus 1)")

    if args.config == "c3":
        batch_main(args)
        return
    if args.config == "c5":
        unrolled_main(args)
        return

    import numpy as np
    import torch

    from unround import native

    image, method, description = workload(args.config)
    log(f"workload {description}; library {native.library_path()}")
    rounded_samples = args.config == "png"

    def solve(options: list[str]):
        if rounded_samples:
            return native.solve_samples(image, 8, options)
        return native.solve(image, options)

    base = ["--method", method, "--tolerance", "0", "--precision", args.precision]
    if args.warmup > 0:
        solve(base + ["--iterations", str(args.warmup), "--record-every", "0"])
    # The timed solve: W untimed iterations, then one pair of events around the K timed
    # ones, nothing between launches.
    os.environ["UNROUND_TIMING"] = "1"
    os.environ["UNROUND_TIMING_WARMUP"] = str(args.warmup)
    torch.cuda.synchronize()
    began = time.perf_counter()
    decoded = solve(base + ["--iterations", str(args.warmup + args.steps), "--record-every", "0"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - began
    times = decoded.kernel_times
    assert times is not None and times.iterations == args.steps, "the library took no timings"
    # The kernels' own durations: events around every launch, in a solve of their own.
    kernel_steps = min(args.steps, 100)
    os.environ["UNROUND_TIMING"] = "2"
    launches = solve(base + ["--iterations", str(args.warmup + kernel_steps), "--record-every", "0"]).kernel_times
    os.environ.pop("UNROUND_TIMING", None)
    os.environ.pop("UNROUND_TIMING_WARMUP", None)
    assert launches is not None and launches.dual_launches == kernel_steps, "the library took no kernel timings"

    channels, rows, columns = decoded.canvas.shape
    samples = channels * rows * columns
    seconds = times.iterations_seconds
    value = samples * args.steps / seconds / 1e6
    # The data term's levels, 2 bytes each: every coefficient's, or every sample's.
    coefficients = samples if rounded_samples else sum(int(np.prod(c.shape)) for c in decoded.coefficients)
    per_sample = 2 * field_bytes(method, args.precision) + 2 * coefficients / samples
    iteration_achieved = per_sample * samples / (seconds / args.steps) / 1e9
    dual_launch = launches.dual_seconds / launches.dual_launches
    primal_launch = launches.primal_seconds / max(launches.primal_launches, 1)
    dual_launch_bytes = dual_bytes(method, args.precision) * samples
    achieved = dual_launch_bytes / dual_launch / 1e9
    log(f"wall {wall:.3f} s for the timed solve ({args.warmup} + {args.steps} iterations + setup + final projection); "
        f"iterations {seconds:.4f} s; per launch (solve of {kernel_steps} iterations timed launch by launch): "
        f"primal kernels {primal_launch * 1e6:.1f} us x {launches.primal_launches}, dual kernel "
        f"{dual_launch * 1e6:.1f} us x {launches.dual_launches}; whole iterations {iteration_achieved:.1f} GB/s "
        f"of {per_sample:.1f} algorithmic bytes per sample")
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
            "height": int(rows),
            "width": int(columns),
            "channels": int(channels),
            "samples_per_iteration": int(samples),
            "precision": args.precision,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "dual_kernel",
            "bytes_per_launch": int(dual_launch_bytes),
            "launch_us": round(dual_launch * 1e6, 2),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("dual_kernel", args.config, args.precision, launches.dual_rows),
            "dual_rows": launches.dual_rows,
            "iteration_achieved": round(iteration_achieved, 1),
            "iteration_frac": round(iteration_achieved / HBM_PEAK_GBS, 4),
        },
        "cpu_baseline": None,
    }
    if rounded_samples:
        # The per-sample proximal step reads x, p_x, p_y and the level and writes x- and x.
        iterate = 8 if args.precision == "double" else 4
        dual = 2 if args.precision == "mixed" else iterate
        primal_bytes = (iterate + 2 * dual + 2 + 2 * iterate) * samples
        primal_achieved = primal_bytes / primal_launch / 1e9
        line["roofline"]["primal"] = {"kernel": "samples_prox_kernel", "bytes_per_launch": int(primal_bytes),
                                      "launch_us": round(primal_launch * 1e6, 2),
                                      "achieved": round(primal_achieved, 1),
                                      "frac": round(primal_achieved / HBM_PEAK_GBS, 4),
                                      "traffic": measured_traffic("samples_prox_kernel", "png", args.precision, 0)}
    if args.config in ("c2", "c4") and args.precision != "double":
        line["accuracy"] = accuracy_against_binary64(args.config, image, decoded, base, args.warmup + args.steps)
    if args.config == "c2" and not args.no_batch:
        line["batch"] = batch_measure(args.precision)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(image, method, samples=rounded_samples)
    print(json.dumps(line), flush=True)


def accuracy_against_binary64(config: str, image, decoded, base: list[str], steps: int) -> dict:
    """The timed solve's result against the same solve in binary64 on the GPU: the largest
    difference of the canvases, PSNR against the uncompressed picture (RGB of a colour one),
    and whether every block of both requantizes to the file's levels."""
    import numpy as np

    from unround import native, synthetic

    options = [value for value in base]
    options[options.index("--precision") + 1] = "double"
    reference = native.solve(image, options + ["--iterations", str(steps), "--record-every", "0"])
    h, w = int(image.height), int(image.width)
    colour = len(image.components) == 3
    original = (synthetic.colour_picture(h, w, 0) if colour else synthetic.grey_picture(h, w, 0)).astype(np.float64)

    def psnr(result):
        picture = native.to_rgb(result.planes) if colour else result.planes[0]
        error = np.mean((np.clip(np.round(picture), 0, 255) - original) ** 2)
        return float(10 * np.log10(255.0**2 / error))

    def exact(result):
        for index, problem in enumerate(result.problems):
            if not np.array_equal(synthetic.requantized(result.canvas[index], image, index, problem.ratio),
                                  np.asarray(image.components[index].coefficients)):
                return False
        return True

    ours, theirs = psnr(decoded), psnr(reference)
    return {
        "reference": "the same solve in binary64 on the GPU",
        "iterations": steps,
        "max_abs_canvas": float(np.max(np.abs(decoded.canvas - reference.canvas))),
        "psnr": round(ours, 5),
        "psnr_binary64": round(theirs, 5),
        "psnr_delta_db": round(ours - theirs, 6),
        "requantizes_exactly": exact(decoded),
        "requantizes_exactly_binary64": exact(reference),
    }


if __name__ == "__main__":
    main()
