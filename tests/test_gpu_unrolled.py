"""The unrolled TGV's forward and hand-written adjoint on the GPU (binary32) against the same
iteration in torch binary64 on the CPU (unrolled_reference.py): the last canvases, the
gradients of alpha0, alpha1 and the weight map against autograd's binary64 ones (relative
1e-3), and against central finite differences of the binary64 loss (h = 1e-6 within 1e-3;
h = 1e-3 within the derivative's range over the interval, the loss being piecewise
smooth); the gradients do not depend on how often the state is kept. Loss: the mean
squared error of the canvases to the uncompressed pictures."""

import numpy as np
import pytest
import torch

import unrolled_reference
from unround import synthetic
from unround.unrolled import UnrolledTGV

pytestmark = pytest.mark.gpu

SIZE, COUNT, ITERATIONS = 64, 2, 20


def pictures():
    originals = [synthetic.grey_picture(SIZE, SIZE, seed) for seed in range(1, COUNT + 1)]
    images = [synthetic.encode_grey(picture, 30) for picture in originals]
    return images, np.stack(originals).astype(np.float64)


def weight_map():
    draw = np.random.default_rng(7)
    return 0.5 + draw.random((SIZE, SIZE))


def gpu_gradients(solve, originals, alpha0=2.0, alpha1=1.0):
    a0 = torch.tensor(alpha0, dtype=torch.float64, requires_grad=True)
    a1 = torch.tensor(alpha1, dtype=torch.float64, requires_grad=True)
    m = torch.tensor(weight_map(), dtype=torch.float32, device="cuda", requires_grad=True)
    canvas = solve(a0, a1, m)
    target = torch.tensor(originals, dtype=torch.float32, device="cuda")
    loss = ((canvas - target) ** 2).mean()
    loss.backward()
    return canvas.detach().cpu().double().numpy(), float(a0.grad), float(a1.grad), m.grad.detach().cpu().double().numpy()


def reference(solve, originals, alpha0=2.0, alpha1=1.0, map_values=None, grad=True):
    start = solve.start_state()
    problems = [solve.problem(k) for k in range(solve.count)]
    a0 = torch.tensor(alpha0, dtype=torch.float64, requires_grad=grad)
    a1 = torch.tensor(alpha1, dtype=torch.float64, requires_grad=grad)
    m = torch.tensor(weight_map() if map_values is None else map_values, dtype=torch.float64, requires_grad=grad)
    x = unrolled_reference.unrolled(start, problems, solve.constants(), a0, a1, m, solve.iterations)
    loss = ((x - torch.tensor(originals)) ** 2).mean()
    if not grad:
        return float(loss)
    loss.backward()
    return x.detach().numpy(), float(a0.grad), float(a1.grad), m.grad.numpy(), float(loss.detach())


def relative(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def test_forward_and_gradients_match_binary64_autograd():
    images, originals = pictures()
    solve = UnrolledTGV(images, ["--method", "tgv"], iterations=ITERATIONS, checkpoint_every=5)
    canvas, g0, g1, gm = gpu_gradients(solve, originals)
    x, r0, r1, rm, _ = reference(solve, originals)
    print(f"canvas max |diff| {np.max(np.abs(canvas - x)):.3g}; alpha0 {g0:.6g} vs {r0:.6g}; "
          f"alpha1 {g1:.6g} vs {r1:.6g}; map rel {relative(gm, rm):.3g}", flush=True)
    assert np.max(np.abs(canvas - x)) < 1e-2
    assert abs(g0 - r0) <= 1e-3 * abs(r0)
    assert abs(g1 - r1) <= 1e-3 * abs(r1)
    assert relative(gm, rm) <= 1e-3


def test_gradients_match_central_differences():
    # The loss is piecewise smooth: each ball's and interval's switch between active and not
    # is a kink, and at h = 1e-3 the central difference averages the derivative over many of
    # them (binary64 autograd itself moves by ~3 % over alpha1 +- 1e-3). So: the difference
    # at h = 1e-6 (binary64, converged) within 1e-3 of the GPU's gradient; and the one at
    # h = 1e-3 within the range the binary64 derivative takes over [a - h, a + h], as the
    # mean value of the derivative over the interval must be.
    images, originals = pictures()
    solve = UnrolledTGV(images, ["--method", "tgv"], iterations=ITERATIONS, checkpoint_every=5)
    _, g0, g1, gm = gpu_gradients(solve, originals)

    def difference(h, name, base):
        return (reference(solve, originals, grad=False, **{name: base + h}) -
                reference(solve, originals, grad=False, **{name: base - h})) / (2 * h)

    for name, base, ours, which in (("alpha0", 2.0, g0, 1), ("alpha1", 1.0, g1, 2)):
        converged = difference(1e-6, name, base)
        wide = difference(1e-3, name, base)
        along = [reference(solve, originals, **{name: base + d})[which] for d in (-1e-3, -5e-4, 0.0, 5e-4, 1e-3)]
        slack = 1e-3 * abs(converged)
        print(f"{name}: gpu {ours:.7g}; fd(1e-6) {converged:.7g}; fd(1e-3) {wide:.7g} in "
              f"[{min(along):.7g}, {max(along):.7g}]", flush=True)
        assert abs(ours - converged) <= 1e-3 * abs(converged)
        assert min(along) - slack <= wide <= max(along) + slack
    direction = np.random.default_rng(11).standard_normal((SIZE, SIZE))
    base = weight_map()
    h = 1e-6
    fdm = (reference(solve, originals, map_values=base + h * direction, grad=False) -
           reference(solve, originals, map_values=base - h * direction, grad=False)) / (2 * h)
    along = float(np.sum(gm * direction))
    print(f"map along a direction: gpu {along:.7g}, fd(1e-6) {fdm:.7g}", flush=True)
    assert abs(along - fdm) <= 1e-3 * abs(fdm)


def test_gradients_do_not_depend_on_the_checkpoints():
    images, originals = pictures()
    results = []
    for every in (1, 7, ITERATIONS):
        solve = UnrolledTGV(images, ["--method", "tgv"], iterations=ITERATIONS, checkpoint_every=every)
        results.append(gpu_gradients(solve, originals))
    for other in results[1:]:
        assert np.array_equal(other[0], results[0][0])
        assert other[1] == results[0][1] and other[2] == results[0][2]
        assert np.array_equal(other[3], results[0][3])


def test_gradients_do_not_depend_on_the_adjoint_s_runs(monkeypatch):
    # The dual step's adjoint walks runs of rows, each starting two rows above its first with
    # the same arithmetic: the map's gradient and alpha1's come out the same bit for bit
    # whatever the runs; alpha0's sums its sensitivities a run at a time (binary32 per run,
    # binary64 over them), so it agrees to rounding.
    images, originals = pictures()
    results = []
    for rows in ("5", "16", "64"):
        monkeypatch.setenv("UNROUND_BACK_ROWS", rows)
        solve = UnrolledTGV(images, ["--method", "tgv"], iterations=ITERATIONS, checkpoint_every=5)
        results.append(gpu_gradients(solve, originals))
        solve.close()
    for other in results[1:]:
        assert np.array_equal(other[0], results[0][0])
        assert other[2] == results[0][2]
        assert np.array_equal(other[3], results[0][3])
        assert abs(other[1] - results[0][1]) <= 1e-6 * abs(results[0][1])


def test_launches_are_timed_where_asked(monkeypatch):
    images, originals = pictures()
    solve = UnrolledTGV(images, ["--method", "tgv"], iterations=ITERATIONS, checkpoint_every=5)
    gpu_gradients(solve, originals)
    assert solve.kernel_times() is None
    monkeypatch.setenv("UNROUND_TIMING", "2")
    timed = UnrolledTGV(images, ["--method", "tgv"], iterations=ITERATIONS, checkpoint_every=5)
    gpu_gradients(timed, originals)
    times = timed.kernel_times()
    # 20 adjoint iterations; the forward's 20 dual steps and the recomputation of the three
    # runs before the last (15).
    assert times["dual_back"][1] == ITERATIONS and times["prox_back"][1] == ITERATIONS
    assert times["dual"][1] == ITERATIONS + 15
    assert all(seconds > 0 for seconds, _ in times.values())
