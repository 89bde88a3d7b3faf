"""The state of the annealing at one temperature (altar/altar/bayesian/CoolingStep.py:17-134).

theta (samples x parameters), prior/data/posterior (samples) and sigma
(parameters x parameters) are FP64 tensors on the GPU.
"""
import torch

from .. import _lib


class CoolingStep:
    def __init__(self, beta, theta, likelihoods, sigma=None):
        self.beta = beta
        self.theta = theta
        self.prior, self.data, self.posterior = likelihoods
        P = theta.shape[1]
        self.sigma = (torch.zeros((P, P), dtype=torch.float64, device=theta.device)
                      if sigma is None else sigma)

    @property
    def samples(self):
        return self.theta.shape[0]

    @property
    def parameters(self):
        return self.theta.shape[1]

    @classmethod
    def alloc(cls, samples, parameters, device="cuda"):
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=device)
        return cls(beta=0.0, theta=z(samples, parameters), likelihoods=(z(samples), z(samples), z(samples)))

    @classmethod
    def start(cls, annealer):
        """CoolingStep.py:53-71: sample the prep distribution, compute likelihoods."""
        model = annealer.model
        step = cls.alloc(model.job.chains, model.parameters, device=model.device)
        model.initializeSample(step=step)
        model.likelihoods(annealer=annealer, step=step)
        return step

    def clone(self):
        return CoolingStep(self.beta, self.theta.clone(),
                           (self.prior.clone(), self.data.clone(), self.posterior.clone()),
                           self.sigma.clone())

    def computePosterior(self):
        """posterior = prior + beta data (CoolingStep.py:103-113)."""
        _lib.call("altar_posterior", _lib.ptr(self.prior), _lib.ptr(self.data), float(self.beta),
                  self.samples, _lib.ptr(self.posterior), _lib.stream())
        return self
