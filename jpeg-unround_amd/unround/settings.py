"""The settings of a reconstruction: the method, the model, and the solvers' options.

The same dataclasses, fields and defaults as the reference's package
(python/src/unround/settings.py:22-148), so that its callers' settings carry over; each
field is an option of the command line, and the library checks their ranges. Where a
value is None, the library takes the model's default (docs/math.md, 6.5).
PdhgOptions.precision is this edition's: "single" (binary32 iterates, the default),
"double" (binary64 iterates) or "mixed" (binary32 iterates, the dual fields stored in
bfloat16).
"""

from dataclasses import dataclass, field
from typing import Literal

__all__ = ["TGV", "TV", "Centres", "DataTerm", "Method", "PdhgOptions", "Precision", "Settings", "SubgradientOptions"]

Method = Literal["mmse", "tv", "tgv", "subgradient"]
Centres = Literal["mmse", "midpoint"]
Precision = Literal["single", "double", "mixed"]


@dataclass(frozen=True, slots=True)
class DataTerm:
    """The options of G (docs/math.md, 4.1): weights, centres, slack and flat blocks."""

    mu: float | None = None
    mu_scale: float = 9.0
    mu_power: float = 0.9
    mu_chroma: float = 0.3
    slack: float = 0.0
    dc_weight: float = 1.0
    centres: Centres = "midpoint"
    power: float = 2.0
    slack_cost: float = 0.0
    flat_weight: float = 0.1
    flat_levels: int = 0
    flat_dc: int | None = 1


@dataclass(frozen=True, slots=True)
class TV:
    """The weight of total variation (4.2) and how it takes several channels (4.4)."""

    alpha: float = 1.0
    channel_weights: tuple[float, ...] | None = None
    coupled: bool = True


@dataclass(frozen=True, slots=True)
class TGV:
    """The weights of second-order total generalized variation (4.3 and 4.4)."""

    alpha1: float = 1.0
    alpha0: float = 2.0
    channel_weights: tuple[float, ...] | None = None
    coupled: bool = True


@dataclass(frozen=True, slots=True)
class PdhgOptions:
    """When the primal-dual method stops, its steps, how often it records, and its arithmetic."""

    iterations: int | None = None
    tolerance: float | None = None
    relative_tolerance: float = 0.0
    partial_tolerance: float = 0.0
    step_ratio: float | None = None
    relaxation: float | None = None
    step_product: float = 0.99
    norm_squared: float | None = None
    scale_with_weight: bool = True
    record_every: int = 10
    partial_radius: float | None = None
    free_radius: float = 255.0
    threads: int = 0
    precision: Precision = "single"


@dataclass(frozen=True, slots=True)
class SubgradientOptions:
    """The subgradient method's options, kept for the round trip of the settings."""

    iterations: int = 50
    record_every: int = 1
    step: float = 0.5
    decay: float = 0.5
    momentum: bool = True


@dataclass(frozen=True, slots=True)
class Settings:
    """The method, the model, and the solvers' options."""

    method: Method = "tgv"
    data: DataTerm | tuple[DataTerm, ...] = field(default_factory=DataTerm)
    tv: TV = field(default_factory=TV)
    tgv: TGV = field(default_factory=TGV)
    pdhg: PdhgOptions = field(default_factory=PdhgOptions)
    subgradient: SubgradientOptions = field(default_factory=SubgradientOptions)
