"""Event bus with the reference's event names (altar/altar/bayesian/Notifier.py:27-52)
and a phase profiler that times them with HIP events on the current stream
(the reference's Profiler.py times wall clock)."""
import collections


class Notifier:
    EVENTS = ("start", "finish", "betaStart", "betaFinish", "walkChainsStart", "walkChainsFinish",
              "resampleStart", "resampleFinish", "samplePosteriorStart", "samplePosteriorFinish",
              "prepareSamplingPDFStart", "prepareSamplingPDFFinish", "chainAdvanceStart",
              "chainAdvanceFinish", "verifyStart", "verifyFinish", "priorStart", "priorFinish",
              "dataStart", "dataFinish", "posteriorStart", "posteriorFinish", "acceptStart",
              "acceptFinish")

    def __init__(self):
        self.monitors = []
        for e in self.EVENTS:
            setattr(self, e, e)

    def subscribe(self, monitor):
        self.monitors.append(monitor)
        return self

    def notify(self, event, controller=None):
        for m in self.monitors:
            m.on(event, controller)


class Profiler:
    """Accumulates the GPU time between matching *Start / *Finish events."""

    def __init__(self):
        import torch
        self.torch = torch
        self.open = {}
        self.pairs = collections.defaultdict(list)

    def on(self, event, controller):
        torch = self.torch
        if not torch.cuda.is_available():
            return
        if event.endswith("Start"):
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            self.open[event[:-5]] = ev
        elif event.endswith("Finish"):
            name = event[:-6]
            start = self.open.pop(name, None)
            if start is not None:
                ev = torch.cuda.Event(enable_timing=True)
                ev.record()
                self.pairs[name].append((start, ev))

    def totals(self):
        """Milliseconds per phase (synchronizes)."""
        self.torch.cuda.synchronize()
        return {k: sum(a.elapsed_time(b) for a, b in v) for k, v in self.pairs.items()}
