"""CATMIP controller: annealer, COV scheduler, Metropolis sampler, cooling step.

Mirrors altar/altar/bayesian/ (Annealer.py, COV.py, Grid.py, Brent.py, Metropolis.py,
CoolingStep.py, Notifier.py) with the sample state resident on the GPU.
"""
from .annealer import Annealer, SequentialAnnealing
from .cooling_step import CoolingStep
from .cov import COV, Brent, Grid
from .metropolis import Metropolis
from .notifier import Notifier, Profiler

__all__ = ["Annealer", "SequentialAnnealing", "CoolingStep", "COV", "Brent", "Grid", "Metropolis",
           "Notifier", "Profiler"]
