"""Prior distributions and contiguous parameter sets.

Mirrors altar/altar/distributions/{Uniform,Gaussian,Base}.py and
altar/altar/models/Contiguous.py.  The arithmetic (support check, log
density, initial draws) runs on the GPU through one ``altar_prior`` handle
per model that concatenates every parameter set's distribution.
"""
from dataclasses import dataclass, field


@dataclass
class Uniform:
    """Uniform on the closed support [a, b] (Uniform.py:44-72 checks
    ``low <= x <= high``); log density -log(b - a) (Base.py:60-79)."""
    support: tuple = (0.0, 1.0)
    kind = 0

    def params(self):
        return float(self.support[0]), float(self.support[1])


@dataclass
class Gaussian:
    """Gaussian(mean, sigma); accepts every value (Gaussian.py:47-55)."""
    mean: float = 0.0
    sigma: float = 1.0
    kind = 1

    def params(self):
        return float(self.mean), float(self.sigma)


@dataclass
class Contiguous:
    """A block of ``count`` consecutive parameters (Contiguous.py:19-127).

    ``prior`` scores and verifies samples; ``prep`` draws the initial sample.
    ``offset`` is set by the model when it lays out its parameter vector.
    """
    count: int = 1
    prior: object = field(default_factory=Uniform)
    prep: object = None
    offset: int = 0

    def initialize(self, model, offset):
        self.offset = offset
        if self.prep is None:
            self.prep = self.prior
        return self.count


def flatten(psets, which="prior"):
    """(kind, a, b) lists over all parameters of an ordered dict of psets."""
    kind, a, b = [], [], []
    for pset in psets.values():
        dist = getattr(pset, which)
        for _ in range(pset.count):
            kind.append(dist.kind)
            lo, hi = dist.params()
            a.append(lo)
            b.append(hi)
    return kind, a, b
