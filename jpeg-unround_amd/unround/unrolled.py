"""Unrolled TGV on the GPU with its hand-written adjoint, as a torch.autograd.Function.

The forward runs a fixed number of the method's iterations from the usual start on grey
pictures of one size, with p's ball of radius alpha1 * m(i, j) at pixel (i, j) and r's of
alpha0, and returns the last x of every picture, (count, height, width) binary32 on the
GPU. The backward takes the loss's gradient on that canvas to the gradients of alpha0,
alpha1 and the map m, through the library's adjoint kernels, recomputing each run of
`checkpoint_every` iterations from the state kept at its start. The weight map has no
counterpart in the reference (one radius per channel, docs/math.md:544-566).

    solve = UnrolledTGV(images, ["--method", "tgv"], iterations=100, checkpoint_every=10)
    canvas = solve(alpha0, alpha1, weight_map)   # alpha0, alpha1: 0-d tensors; map (H, W)
    loss = ((canvas - originals) ** 2).mean()
    loss.backward()

A handle keeps the checkpoints of its last forward: call backward once per forward, before
the next forward on the same handle.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np
import torch

from . import jpegio, native

_DOUBLES = ctypes.POINTER(ctypes.c_double)

# The unrolled solve's entry points (include/unround_ext.h), in native.declare's table form.
_ENTRY_POINTS = """
unround_unrolled_new        i32   input* size handle u32 u32 handle* message
unround_unrolled_free       void  handle
unround_unrolled_shape      i32   handle u64* u64* u64*
unround_unrolled_constants  i32   handle f64*
unround_unrolled_start      i32   handle f64* message
unround_unrolled_problem    i32   handle u64 f64*
unround_unrolled_forward    i32   handle address f64 f64 address address message
unround_unrolled_backward   i32   handle address address f64* address message
unround_unrolled_kernel_times  i32  handle f64* u64
"""

_declared_library: ctypes.CDLL | None = None


def _declared() -> ctypes.CDLL:
    global _declared_library
    if _declared_library is None:
        library = native._lib()
        native.declare(library, _ENTRY_POINTS)
        _declared_library = library
    return _declared_library


def _call(function, *arguments: object) -> None:
    """A call of an entry point whose last arguments are the message buffer and its size."""
    room = ctypes.create_string_buffer(native.MESSAGE_BYTES)
    status = int(function(*arguments, room, native.MESSAGE_BYTES))
    if status != native.Status.OK:
        raise native.status_error(status, room.value.decode("utf-8", "replace"))


class UnrolledTGV:
    """Unrolled TGV over grey pictures of one size (see the module's docstring)."""

    def __init__(self, images: Sequence[jpegio.Image], settings: native.Settings | list[str] | None = None, *,
                 iterations: int = 100, checkpoint_every: int = 10) -> None:
        library = _declared()
        self._library = library
        self._settings = native.SettingsHandle.of(settings)
        marshalled = [native.Marshalled(image) for image in images]
        count = len(marshalled)
        inputs = (native.CInput * max(count, 1))(*(m.input for m in marshalled))
        handle = ctypes.c_void_p()
        _call(library.unround_unrolled_new, inputs, count, self._settings.handle, iterations, checkpoint_every,
              ctypes.byref(handle))
        self._handle = handle
        shape = [ctypes.c_uint64() for _ in range(3)]
        library.unround_unrolled_shape(handle, *(ctypes.byref(value) for value in shape))
        self.count, self.height, self.width = (int(value.value) for value in shape)
        self.iterations = iterations
        self.checkpoint_every = checkpoint_every

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None and self._handle.value is not None:
            self._library.unround_unrolled_free(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self) -> None:
        self.close()

    def constants(self) -> dict[str, float]:
        """tau, sigma, the channel's gamma and the relaxation rho of the pictures' plan."""
        values = (ctypes.c_double * 4)()
        self._library.unround_unrolled_constants(self._handle, values)
        return dict(zip(("tau", "sigma", "gamma", "rho"), (float(v) for v in values)))

    def start_state(self) -> np.ndarray:
        """The start, (count, 8, height, width): x, w_x, w_y, p_x, p_y, r_xx, r_yy, r_xy."""
        out = np.zeros((self.count, 8, self.height, self.width))
        _call(self._library.unround_unrolled_start, self._handle, out.ctypes.data_as(_DOUBLES))
        return out

    def problem(self, frame: int) -> np.ndarray:
        """Picture `frame`'s coefficients, (rows, columns, 64, 4): lower, upper, centre, weight."""
        rows, columns = self.height // 8, self.width // 8
        out = np.zeros((rows, columns, 64, 4))
        status = self._library.unround_unrolled_problem(self._handle, frame, out.ctypes.data_as(_DOUBLES))
        if status != native.Status.OK:
            raise native.UnroundError(native.Status.ARGUMENT, f"no picture {frame}")
        return out

    def kernel_times(self) -> dict[str, tuple[float, int]] | None:
        """Seconds and launches since the last forward began of the dual step's adjoint
        ("dual_back"), the prox's adjoint ("prox_back") and the forward dual step ("dual"),
        where the handle was made with UNROUND_TIMING=2 in the environment; else None."""
        values = (ctypes.c_double * 6)()
        if not self._library.unround_unrolled_kernel_times(self._handle, values, 6):
            return None
        names = ("dual_back", "prox_back", "dual")
        return {name: (float(values[2 * k]), int(values[2 * k + 1])) for k, name in enumerate(names)}

    def __call__(self, alpha0: torch.Tensor, alpha1: torch.Tensor, weight_map: torch.Tensor) -> torch.Tensor:
        return _Unroll.apply(alpha0, alpha1, weight_map, self)

    def _forward(self, weight_map: torch.Tensor, alpha0: float, alpha1: float) -> torch.Tensor:
        canvas = torch.empty((self.count, self.height, self.width), device=weight_map.device, dtype=torch.float32)
        stream = torch.cuda.current_stream(weight_map.device).cuda_stream
        _call(self._library.unround_unrolled_forward, self._handle, weight_map.data_ptr(), alpha0, alpha1,
              canvas.data_ptr(), stream)
        return canvas

    def _backward(self, grad_canvas: torch.Tensor) -> tuple[float, float, torch.Tensor]:
        grad_map = torch.empty((self.height, self.width), device=grad_canvas.device, dtype=torch.float32)
        alphas = (ctypes.c_double * 2)()
        stream = torch.cuda.current_stream(grad_canvas.device).cuda_stream
        _call(self._library.unround_unrolled_backward, self._handle, grad_canvas.data_ptr(), grad_map.data_ptr(),
              alphas, stream)
        return float(alphas[0]), float(alphas[1]), grad_map


class _Unroll(torch.autograd.Function):
    """The library's forward and adjoint behind autograd: alpha0, alpha1 (0-d tensors) and the
    weight map (height, width) to the canvases (count, height, width)."""

    @staticmethod
    def forward(ctx, alpha0: torch.Tensor, alpha1: torch.Tensor, weight_map: torch.Tensor,
                solve: UnrolledTGV) -> torch.Tensor:
        if not weight_map.is_cuda:
            raise ValueError("the weight map lives on the GPU")
        if tuple(weight_map.shape) != (solve.height, solve.width):
            raise ValueError(f"a weight map of ({solve.height}, {solve.width}), not {tuple(weight_map.shape)}")
        map32 = weight_map.detach().to(torch.float32).contiguous()
        ctx.solve = solve
        ctx.map32 = map32
        ctx.kinds = (alpha0, alpha1, weight_map.dtype)
        return solve._forward(map32, float(alpha0), float(alpha1))

    @staticmethod
    def backward(ctx, grad_canvas: torch.Tensor):
        alpha0, alpha1, map_dtype = ctx.kinds
        grad = grad_canvas.detach().to(torch.float32).contiguous()
        g0, g1, grad_map = ctx.solve._backward(grad)

        def like(value: float, reference: torch.Tensor) -> torch.Tensor:
            if isinstance(reference, torch.Tensor):
                return torch.tensor(value, dtype=reference.dtype, device=reference.device)
            return torch.tensor(value)

        return like(g0, alpha0), like(g1, alpha1), grad_map.to(map_dtype), None
