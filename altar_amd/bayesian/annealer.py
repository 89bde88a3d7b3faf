"""The CATMIP annealer (altar/altar/bayesian/Annealer.py:78-135) and its
sequential worker (AnnealingMethod.py, SequentialAnnealing.py)."""
import time

from .cooling_step import CoolingStep
from .cov import COV
from .metropolis import Metropolis
from .notifier import Notifier


class SequentialAnnealing:
    """AnnealingMethod.py:13-155 on one GPU."""

    def __init__(self, annealer):
        self.annealer = annealer
        self.step = None
        self.iteration = 0

    @property
    def beta(self):
        return self.step.beta

    def start(self, annealer):
        self.step = CoolingStep.start(annealer)
        self.iteration = 0
        return self

    def top(self, annealer):
        annealer.model.top(step=self.step)
        return self

    def cool(self, annealer):
        self.iteration += 1
        annealer.scheduler.update(step=self.step, annealer=annealer)
        return self

    def walk(self, annealer):
        return annealer.sampler.samplePosterior(annealer=annealer, step=self.step)

    def resample(self, annealer, statistics):
        annealer.sampler.resample(annealer=annealer, statistics=statistics)
        return self

    def bottom(self, annealer):
        annealer.model.bottom(step=self.step)
        return self

    def finish(self, annealer):
        if annealer.archiver is not None:
            annealer.archiver.record(step=self.step)
        return self


class Annealer:
    def __init__(self, sampler=None, scheduler=None, dispatcher=None, archiver=None, seed=0):
        self.sampler = sampler or Metropolis()
        self.scheduler = scheduler or COV()
        self.dispatcher = dispatcher or Notifier()
        self.archiver = archiver
        self.seed = seed
        self.model = None
        self.worker = SequentialAnnealing(self)
        self.history = []

    def initialize(self, job=None):
        self.sampler.initialize(job=job)
        self.scheduler.initialize()
        if job is not None:
            self.seed = job.seed
        return self

    def posterior(self, model):
        """Annealer.py:78-135: iterate until beta + tolerance >= 1."""
        self.model = model
        tolerance = model.job.tolerance
        worker = self.worker
        d = self.dispatcher
        d.notify(d.start, self)
        worker.start(annealer=self)
        while worker.beta + tolerance < 1:
            d.notify(d.betaStart, self)
            worker.top(annealer=self)
            worker.cool(annealer=self)
            d.notify(d.walkChainsStart, self)
            stats = worker.walk(annealer=self)
            d.notify(d.walkChainsFinish, self)
            d.notify(d.resampleStart, self)
            worker.resample(annealer=self, statistics=stats)
            d.notify(d.resampleFinish, self)
            worker.bottom(annealer=self)
            d.notify(d.betaFinish, self)
            self.history.append({"iteration": worker.iteration, "beta": worker.beta,
                                 "cov": self.scheduler.cov, "accepted": stats[0], "rejected": stats[1],
                                 "unlikely": stats[2], "scaling": self.sampler.scaling,
                                 "time": time.time()})
        worker.finish(annealer=self)
        d.notify(d.finish, self)
        self.model = None
        return 0
