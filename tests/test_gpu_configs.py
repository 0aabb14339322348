"""The benchmarked workloads at their own sizes.

C2 (2048 x 2048 RGB 4:2:0 q75, TGV) and C4 (4096 x 4096 grey q40, TGV): 20 iterations on
the GPU against the oracle (threaded over row bands, binary64) with the parity criteria of
the smaller cases: 8-bit outputs within 1 level on at most 0.1 % of the samples, PSNR
within 0.01 dB, every coefficient requantizing to its level, records within 1e-4. At
these sizes the dual step runs in 32- and 64-row runs, which no smaller case reaches.

C3 (the 64 mixed pictures, sides 96..1920): the batch at full size, by properties: every
picture's coefficients inside their intervals and requantizing to their levels, stopping
where the picture's own solve stops with the same canvas; a run loose enough for most
pictures to stop by their own relative gap. The 4 smallest against the oracle.

C5 (16 grey 512 x 512 q30, 100 unrolled iterations, the state kept every 10): the
gradients the same whatever the checkpoints; the central difference of the loss (the GPU's
canvases, the loss summed in binary64) at h = 1e-3 within the range the GPU's derivative
takes over [a - h, a + h]; the forward with m = 1 against the oracle's last iterate."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import helpers
from unround import native, synthetic
from unround.unrolled import UnrolledTGV

pytestmark = pytest.mark.gpu


def test_c2_full_size_against_the_oracle():
    picture = synthetic.colour_picture(2048, 2048, 0)
    image = synthetic.encode_colour420(picture, 75)
    helpers.parity_check(image, picture, ["--method", "tgv", "--iterations", "20", "--tolerance", "0"])


def test_c4_full_size_against_the_oracle():
    picture = synthetic.grey_picture(4096, 4096, 0)
    image = synthetic.encode_grey(picture, 40)
    helpers.parity_check(image, picture, ["--method", "tgv", "--iterations", "20", "--tolerance", "0"])


def test_c4_full_size_bf16_duals_requantize():
    # BF16 p and r (8-bit significands): the distance to binary64 is reported; the
    # requantization stays exact and every coefficient inside its interval.
    picture = synthetic.grey_picture(4096, 4096, 0)
    image = synthetic.encode_grey(picture, 40)
    options = ["--method", "tgv", "--iterations", "20", "--tolerance", "0"]
    print("C4 mixed, 20 iterations", flush=True)
    expected = helpers.oracle_solve(image, options)
    decoded = native.solve(image, options + ["--precision", "mixed"])
    worst = float(np.max(np.abs(decoded.planes[0] - expected.canvas[0])))
    ours, theirs = helpers.psnr(decoded.planes[0], picture), helpers.psnr(expected.canvas[0], picture)
    print(f"  max |diff| {worst:.3g}, PSNR {ours:.4f} vs {theirs:.4f}", flush=True)
    assert worst < 0.5
    assert abs(ours - theirs) < 0.01
    problem = decoded.problems[0]
    assert np.all(decoded.coefficients[0] > problem.lower) and np.all(decoded.coefficients[0] < problem.upper)
    levels = helpers.requantized(decoded.canvas[0], image, 0, (1, 1), 512, 512)
    assert np.array_equal(levels, helpers.kept_levels(image, 0, 512, 512))


def inside_and_requantizing(decoded, image, label):
    for c, problem in enumerate(decoded.problems):
        rows, columns = decoded.coefficients[c].shape[:2]
        assert np.all(decoded.coefficients[c] > problem.lower), f"{label}: component {c} below"
        assert np.all(decoded.coefficients[c] < problem.upper), f"{label}: component {c} above"
        levels = helpers.requantized(decoded.canvas[c], image, c, problem.ratio, rows, columns)
        assert np.array_equal(levels, helpers.kept_levels(image, c, rows, columns)), f"{label}: component {c}"


def test_c3_batch_at_full_size():
    images, originals, kinds = synthetic.batch_pictures()
    options = ["--method", "tgv", "--tolerance", "0", "--relative-tolerance", "1e-3", "--record-every", "10",
               "--iterations", "3000"]
    print(f"C3: {len(images)} pictures, {' '.join(options)}", flush=True)
    together, latencies = native.solve_batch(images, options)
    assert len(together) == len(images) and len(latencies) == len(images)
    stops = [d.result.stop for d in together]
    by_gap = [k for k, stop in enumerate(stops) if stop == native.Stop.CONVERGED]
    iterations = [d.result.iterations for d in together]
    print(f"  stopped by the gap: {len(by_gap)} of {len(images)}; iterations {min(iterations)}..{max(iterations)}",
          flush=True)
    assert len(by_gap) >= len(images) // 4
    for k, (decoded, image) in enumerate(zip(together, images)):
        inside_and_requantizing(decoded, image, f"picture {k} ({image.height}x{image.width} {kinds[k]})")
        assert decoded.result.iterations % 10 == 0
    # Eight pictures alone: the smallest three, the largest two, three that stopped late.
    sizes = np.argsort([image.height * image.width for image in images])
    late = sorted(by_gap, key=lambda k: -iterations[k])[:3]
    chosen = list(dict.fromkeys([*sizes[:3].tolist(), *sizes[-2:].tolist(), *late]))
    for k in chosen:
        print(f"  alone: picture {k} {images[k].height}x{images[k].width}", flush=True)
        alone = native.solve(images[k], options)
        assert alone.result.iterations == together[k].result.iterations, k
        assert alone.result.stop == together[k].result.stop, k
        assert np.array_equal(alone.canvas, together[k].canvas), k
        assert np.array_equal(alone.result.history.primal, together[k].result.history.primal), k


def test_c3_smallest_four_against_the_oracle():
    images, originals, _ = synthetic.batch_pictures()
    smallest = np.argsort([image.height * image.width for image in images])[:4].tolist()
    options = ["--method", "tgv", "--iterations", "100", "--tolerance", "0"]
    print(f"C3 smallest four: {smallest}", flush=True)
    together, _ = native.solve_batch([images[k] for k in smallest], options)
    for decoded, k in zip(together, smallest):
        print(f"  picture {k}: {images[k].height}x{images[k].width}", flush=True)
        expected = helpers.oracle_solve(images[k], options)
        helpers.parity_compare(decoded, expected, images[k], originals[k], options)


C5_COUNT, C5_SIZE, C5_ITERATIONS, C5_EVERY = 16, 512, 100, 10


def c5_pictures():
    originals = [synthetic.grey_picture(C5_SIZE, C5_SIZE, seed) for seed in range(1, C5_COUNT + 1)]
    images = [synthetic.encode_grey(picture, 30) for picture in originals]
    return images, np.stack(originals).astype(np.float64)


def c5_run(solve, target, alpha0=2.0, alpha1=1.0, grad=True):
    """The loss (mean squared error, summed in binary64) and, if asked, its gradients."""
    a0 = torch.tensor(alpha0, dtype=torch.float64, requires_grad=grad)
    a1 = torch.tensor(alpha1, dtype=torch.float64, requires_grad=grad)
    m = torch.ones((C5_SIZE, C5_SIZE), dtype=torch.float32, device="cuda", requires_grad=grad)
    canvas = solve(a0, a1, m)
    if not grad:
        return float(((canvas.double() - target) ** 2).mean())
    loss = ((canvas.double() - target) ** 2).mean()
    loss.backward()
    return float(loss.detach()), float(a0.grad), float(a1.grad), m.grad.detach().cpu().double().numpy()


def test_c5_gradients_do_not_depend_on_the_checkpoints():
    images, originals = c5_pictures()
    target = torch.tensor(originals, device="cuda")
    results = []
    for every in (C5_EVERY, 7, C5_ITERATIONS):
        print(f"C5: checkpoint every {every}", flush=True)
        solve = UnrolledTGV(images, ["--method", "tgv"], iterations=C5_ITERATIONS, checkpoint_every=every)
        results.append(c5_run(solve, target))
        solve.close()
    for other in results[1:]:
        assert other[:3] == results[0][:3]
        assert np.array_equal(other[3], results[0][3])


def test_c5_central_differences():
    images, originals = c5_pictures()
    target = torch.tensor(originals, device="cuda")
    solve = UnrolledTGV(images, ["--method", "tgv"], iterations=C5_ITERATIONS, checkpoint_every=C5_EVERY)
    _, g0, g1, _ = c5_run(solve, target)
    h = 1e-3
    for name, base, ours, which in (("alpha0", 2.0, g0, 1), ("alpha1", 1.0, g1, 2)):
        plus = c5_run(solve, target, grad=False, **{name: base + h})
        minus = c5_run(solve, target, grad=False, **{name: base - h})
        difference = (plus - minus) / (2 * h)
        along = [c5_run(solve, target, **{name: base + d})[which] for d in (-h, -h / 2, 0.0, h / 2, h)]
        slack = 1e-3 * max(abs(v) for v in along)
        relative = abs(difference - ours) / abs(ours)
        print(f"C5 {name}: gpu {ours:.7g}; fd(1e-3) {difference:.7g} (relative {relative:.3g}) in "
              f"[{min(along):.7g}, {max(along):.7g}]", flush=True)
        assert min(along) - slack <= difference <= max(along) + slack


def test_c5_forward_with_unit_map_is_the_method():
    # With m = 1 and the method's alphas the unrolled forward is the method's iteration: its
    # canvas after a few iterations is the oracle's last iterate x of each picture.
    images, _ = c5_pictures()
    iterations = 10
    solve = UnrolledTGV(images, ["--method", "tgv"], iterations=iterations, checkpoint_every=5)
    print(f"C5 forward, {iterations} iterations, against the oracle", flush=True)
    constants = solve.constants()
    with torch.no_grad():
        canvas = solve(torch.tensor(2.0, dtype=torch.float64), torch.tensor(1.0, dtype=torch.float64),
                       torch.ones((C5_SIZE, C5_SIZE), dtype=torch.float32, device="cuda")).cpu().double().numpy()
    start = solve.start_state()
    worst = 0.0
    for k, image in enumerate(images):
        expected = helpers.oracle_solve(image, ["--method", "tgv", "--iterations", str(iterations), "--tolerance", "0",
                                                "--record-every", "0"])
        assert abs(constants["tau"] - expected.tau) <= 1e-12 * expected.tau
        assert abs(constants["sigma"] - expected.sigma) <= 1e-12 * expected.sigma
        oracle_start, _ = helpers.oracle_unrolled_inputs(expected)
        assert np.max(np.abs(start[k, :3] - oracle_start[:3])) < 1e-3, k
        assert not np.any(start[k, 3:])
        worst = max(worst, float(np.max(np.abs(canvas[k] - expected.iterate[0]))))
    print(f"  max |diff| {worst:.3g}", flush=True)
    assert worst < 1e-2
