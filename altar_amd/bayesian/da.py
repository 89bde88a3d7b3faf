"""Two-stage delayed-acceptance Metropolis with a learned surrogate.

Christen and Fox (2005): stage 1 screens the proposal with the surrogate
posterior pi*, stage 2 corrects with the exact posterior pi,

    alpha1 = min(1, pi*(theta') / pi*(theta))
    alpha2 = min(1, pi(theta') pi*(theta) / (pi(theta) pi*(theta'))),

so the chain keeps pi as its stationary distribution whatever the surrogate's
error; the exact forward model runs only for the chains that pass stage 1
(compacted into a dense batch).  No reference counterpart: this extends
altar/altar/bayesian/Metropolis.py:114-233.

With ``substeps`` > 1 each exact step instead follows a surrogate
trajectory (Liu 2001, "Monte Carlo Strategies in Scientific Computing",
the surrogate transition method): ``substeps`` Metropolis steps of pi* lead
from the state theta to an end point phi, which the exact model judges with
the same ratio,

    alpha = min(1, pi(phi) pi*(theta) / (pi(theta) pi*(phi))).

A Metropolis kernel is reversible with respect to its own target, so its
k-fold composition is too, and the ratio keeps pi invariant.  The exact
model runs only for the chains whose trajectory moved; an accepted move has
travelled up to ``substeps`` surrogate steps.

Schedule: while beta < ``train_beta`` the sampler runs plain Metropolis steps
and stores the whitened residuals of every exact evaluation (the training
fields).  When beta reaches ``train_beta`` (or the store is full) the
surrogate is fitted; afterwards each stage validates the surrogate on fresh
exact fields of the current chains and falls back to the plain path when the
field RMSE exceeds ``max_rmse`` metres.  With ``retrain`` (the default) such
a plain stage also replaces the stored fields by its own exact evaluations,
which come from the current, narrower posterior, and the surrogate is
refitted on them before the next stage's validation (retraining on drift).
"""
import numpy as np
import torch

from .. import _lib
from .cooling_step import CoolingStep
from .metropolis import Metropolis


class DelayedAcceptance(Metropolis):
    def __init__(self, surrogate, train_beta=0.1, max_rmse=2e-3, fit_kwargs=None, validate_chains=1024,
                 retrain=True, substeps=1, substep_scale=1.0, substep_target=None, fused=True, **kwds):
        super().__init__(**kwds)
        if not 1 <= substeps <= _lib.TRAJ_MAX_SUBSTEPS:
            raise ValueError(f"substeps must lie in [1, {_lib.TRAJ_MAX_SUBSTEPS}]")
        self.surrogate = surrogate
        self.substeps = substeps
        self.substep_scale = substep_scale   # sub-step proposal = substep_scale * (L xi)
        self.fused = fused                   # whole trajectories in one launch where supported
        # with a target sub-step acceptance rate, each stage first tunes
        # substep_scale by pilot trajectories (which leave the chains as they are)
        self.substep_target = substep_target
        self.substep_counts = [0, 0, 0]   # surrogate sub-steps: accepted, rejected by the model, unlikely
        self.retrain = retrain
        self._refit = False
        self.refits = 0
        self.train_beta = train_beta
        self.max_rmse = max_rmse
        self.fit_kwargs = fit_kwargs or {}
        self.validate_chains = validate_chains
        self.mode_history = []
        self.screened = 0
        self.exact_evals = 0

    # -- plain steps that also collect training fields
    def _collecting(self, step):
        return (not self.surrogate.trained and step.beta < self.train_beta and self.surrogate.space() > 0
                and step.samples % 128 == 0)

    def walkChains(self, annealer, step):
        sur = self.surrogate
        if not sur.trained and (step.beta >= self.train_beta or sur.space() < step.samples):
            if sur.count >= 1024:
                sur.fit(**self.fit_kwargs)
        elif self._refit and sur.count >= 1024:
            # fields of the previous (plain) stage: refit on the current posterior
            sur.fit(**self.fit_kwargs)
            self._refit = False
            self.refits += 1
        use_da = sur.trained and self._validate(annealer, step)
        self.mode_history.append({"beta": step.beta, "mode": "da" if use_da else "plain",
                                  "rmse": sur.validation_rmse})
        if use_da:
            if self.substeps > 1:
                return self._walk_trajectory(annealer, step)
            return self._walk_da(annealer, step)
        if self._collecting(step):
            return self._walk_plain(annealer, step, collect=True)
        if sur.trained and self.retrain and step.samples % 128 == 0:
            # the surrogate has drifted from the posterior: keep this stage's
            # exact fields (replacing the stored ones) for a refit
            sur.count = 0
            self._refit = True
            return self._walk_plain(annealer, step, collect=True)
        stats = super().walkChains(annealer, step)
        self.exact_evals += step.samples * self.steps
        return stats

    def _validate(self, annealer, step):
        """Field RMSE of the surrogate on the current chains (exact fields)."""
        sur = self.surrogate
        n = min(self.validate_chains, step.samples) // 128 * 128
        if n == 0:
            return sur.validation_rmse is not None and sur.validation_rmse <= self.max_rmse
        theta = step.theta[:n].contiguous()
        llk = torch.empty(n, dtype=torch.float64, device=theta.device)
        wr = torch.empty((n, sur.Npad), dtype=torch.float32, device=theta.device)
        _lib.call("altar_source_data_llk_fields", self.model_source(annealer).handle, _lib.ptr(theta), n,
                  _lib.ptr(llk), _lib.PRECISIONS[sur.fields_precision(annealer.model.precision)], _lib.ptr(wr),
                  sur.Npad, _lib.stream())
        rmse = sur.field_rmse_device(theta, wr)
        sur.validation_rmse = rmse
        return rmse <= self.max_rmse

    @staticmethod
    def model_source(annealer):
        return annealer.model.source

    def _walk_plain(self, annealer, step, collect):
        """Metropolis.walkChains with the candidate's exact llk stored as training data."""
        model = annealer.model
        strategy = model.strategy
        sur = self.surrogate
        original = strategy.dataLikelihood

        def collecting(model, step):
            theta = model.restrict(step.theta)
            if sur.collect_llk(theta, step.data, model.precision) == 0:
                model.source.data_llk(theta, step.data, model.precision)
            return strategy

        strategy.dataLikelihood = collecting
        try:
            stats = super().walkChains(annealer, step)
        finally:
            strategy.dataLikelihood = original
        self.exact_evals += step.samples * self.steps
        return stats

    def _workspace(self, C, P, dev):
        f64 = lambda *s: torch.empty(s, dtype=torch.float64, device=dev)
        ws = getattr(self, "_dws", None)
        if ws is None or ws["cand"].shape != (C, P):
            ws = {"cand": f64(C, P), "xi": f64(C, P), "cprior": f64(C), "cdata_s": f64(C), "cpost_s": f64(C),
                  "post_s": f64(C), "llk_s": f64(C), "packed": f64(C, P), "cdata": f64(C),
                  "phi": f64(C, P), "prior_phi": f64(C), "post_phi": f64(C),
                  "mask": torch.empty(C, dtype=torch.int32, device=dev),
                  "index": torch.empty(2 * C, dtype=torch.int32, device=dev),
                  "count": torch.zeros(1, dtype=torch.int32, device=dev),
                  "z": torch.empty((C, self.surrogate.K), dtype=torch.float32, device=dev)}
            self._dws = ws
        return ws

    def _walk_da(self, annealer, step):
        model = annealer.model
        sur = self.surrogate
        C, P = step.samples, step.parameters
        dev = step.theta.device
        seed = annealer.seed
        stage = annealer.worker.iteration
        ws = self._workspace(C, P, dev)
        s = _lib.stream()
        beta = float(step.beta)
        # surrogate posterior of the current states
        sur.data_llk(step.theta, ws["llk_s"], ws["z"])
        _lib.call("altar_posterior", _lib.ptr(step.prior), _lib.ptr(ws["llk_s"]), beta, C, _lib.ptr(ws["post_s"]), s)
        counters = torch.zeros(4, dtype=torch.int64, device=dev)
        candidate = CoolingStep(beta, ws["cand"], (ws["cprior"], ws["cdata_s"], ws["cpost_s"]), sigma=step.sigma)
        d = annealer.dispatcher
        for k in range(self.steps):
            d.notify(d.chainAdvanceStart, annealer)
            _lib.call("altar_mh_propose", _lib.ptr(step.theta), C, P, _lib.ptr(self.sigma_chol), seed, stage, k,
                      _lib.ptr(ws["xi"]), _lib.ptr(ws["cand"]), s)
            ws["mask"].zero_()
            model.verify(candidate, ws["mask"])
            _lib.call("altar_mh_restore", _lib.ptr(ws["cand"]), _lib.ptr(step.theta), _lib.ptr(ws["mask"]), C, P, s)
            model.priorLikelihood(candidate)
            sur.data_llk(ws["cand"], ws["cdata_s"], ws["z"])
            _lib.call("altar_da_screen", _lib.ptr(ws["cprior"]), _lib.ptr(ws["cdata_s"]), _lib.ptr(ws["post_s"]),
                      _lib.ptr(ws["mask"]), beta, C, seed, stage, k, _lib.ptr(ws["cpost_s"]), _lib.ptr(ws["index"]),
                      _lib.ptr(ws["count"]), _lib.ptr(counters), s)
            n = int(ws["count"].item())
            if n > 0:
                _lib.call("altar_da_pack", _lib.ptr(ws["cand"]), _lib.ptr(ws["index"]), n, P, _lib.ptr(ws["packed"]),
                          s)
                packed = ws["packed"][:n]
                model.source.data_llk(model.restrict(packed), ws["cdata"][:n], model.precision)
                _lib.call("altar_da_accept", _lib.ptr(step.theta), _lib.ptr(step.prior), _lib.ptr(step.data),
                          _lib.ptr(step.posterior), _lib.ptr(ws["post_s"]), _lib.ptr(ws["cand"]),
                          _lib.ptr(ws["cprior"]), _lib.ptr(ws["cdata"]), _lib.ptr(ws["cpost_s"]),
                          _lib.ptr(ws["index"]), n, beta, P, seed, stage, k, None, _lib.ptr(counters), s)
            self.exact_evals += n
            d.notify(d.chainAdvanceFinish, annealer)
        acc, rej, unl2, scr = (int(v) for v in counters.cpu().tolist())
        self.screened += scr
        return acc, rej, unl2 + scr

    def fused_supported(self, model, P):
        """Whether altar_da_trajectory_fused takes this model and network:
        one model over the whole sample, at most 32 parameters, a network of
        hidden width 256 with at most 64 inputs and outputs."""
        m = self.surrogate.mlp
        return bool(self.fused and P <= 32 and m.nhid == 256 and m.nin <= 64 and m.nout <= 64
                    and 2 <= m.nlayers <= 8 and getattr(model, "offset", 0) == 0 and model.parameters == P
                    and hasattr(model, "_prior") and hasattr(model, "source"))

    PILOT_STEP = 1 << 21     # step index of the pilot trajectories' streams (the walk's steps stay below)

    def tune_substep_scale(self, annealer, step, ws, rounds=4, substeps=16, tolerance=0.05):
        """Scale the sub-step proposal so that the surrogate sub-steps accept
        about ``substep_target`` of their proposals.  Each round runs one
        fused trajectory of ``substeps`` sub-steps from the current states into
        the workspace (the chains do not move) and multiplies the scale by
        sqrt(rate / target), within [1/2, 2].  The scale is a function of the
        whole population, like the proposal covariance, and stays fixed
        during the walk."""
        model, sur = annealer.model, self.surrogate
        C, P = step.samples, step.parameters
        p = _lib.ptr
        cnt = torch.zeros(3, dtype=torch.int64, device=step.theta.device)
        target = float(self.substep_target)
        for r in range(rounds):
            L = (self.sigma_chol * self.substep_scale).contiguous()
            cnt.zero_()
            _lib.call("altar_da_trajectory_fused", sur.mlp.handle, model._prior.handle, model.source.handle,
                      p(step.theta), p(step.prior), p(ws["post_s"]), p(ws["phi"]), p(ws["prior_phi"]),
                      p(ws["post_phi"]), p(L), p(sur.e_d), p(sur.s_d), sur.K, float(sur.cperp),
                      float(model.normalization), float(step.beta), C, P, annealer.seed,
                      annealer.worker.iteration, self.PILOT_STEP + r, substeps, p(ws["index"]), p(cnt),
                      _lib.stream())
            a, rj, u = (int(v) for v in cnt.cpu().tolist())
            rate = a / max(a + rj + u, 1)
            if abs(rate - target) < tolerance:
                break
            self.substep_scale *= float(np.clip(np.sqrt(max(rate, 1e-3) / target), 0.5, 2.0))
        return self.substep_scale

    def _walk_trajectory(self, annealer, step):
        """Exact steps along surrogate trajectories of ``substeps`` sub-steps.

        Returns (accepted, not moved or rejected, unlikely at the exact
        stage) per exact step, so the proposal scaling follows the rate at
        which chains change state.  With ``fused`` (and a model and network
        the fused kernel supports) each trajectory is one launch
        (altar_da_trajectory_fused); otherwise every sub-step is a sequence
        of separate calls.  Both give the same trajectories."""
        model = annealer.model
        sur = self.surrogate
        C, P = step.samples, step.parameters
        dev = step.theta.device
        seed = annealer.seed
        stage = annealer.worker.iteration
        ws = self._workspace(C, P, dev)
        s = _lib.stream()
        beta = float(step.beta)
        p = _lib.ptr
        fused = self.fused_supported(model, P)
        if not fused:
            sur.data_llk(step.theta, ws["llk_s"], ws["z"])
            _lib.call("altar_posterior", p(step.prior), p(ws["llk_s"]), beta, C, p(ws["post_s"]), s)
        counters = torch.zeros(4, dtype=torch.int64, device=dev)
        sub_counters = torch.zeros(3, dtype=torch.int64, device=dev)
        candidate = CoolingStep(beta, ws["cand"], (ws["cprior"], ws["cdata_s"], ws["cpost_s"]), sigma=step.sigma)
        if fused and self.substep_target is not None:
            self.tune_substep_scale(annealer, step, ws)
        L = self.sigma_chol if self.substep_scale == 1.0 else (self.sigma_chol * self.substep_scale).contiguous()
        d = annealer.dispatcher
        moved_total = 0
        for k in range(self.steps):
            d.notify(d.chainAdvanceStart, annealer)
            if fused:
                _lib.call("altar_da_trajectory_fused", sur.mlp.handle, model._prior.handle, model.source.handle,
                          p(step.theta), p(step.prior), p(ws["post_s"]), p(ws["phi"]), p(ws["prior_phi"]),
                          p(ws["post_phi"]), p(L), p(sur.e_d), p(sur.s_d), sur.K, float(sur.cperp),
                          float(model.normalization), beta, C, P, seed, stage, k, self.substeps, p(ws["index"]),
                          p(sub_counters), s)
            else:
                ws["phi"].copy_(step.theta)
                ws["prior_phi"].copy_(step.prior)
                ws["post_phi"].copy_(ws["post_s"])
            for j in range(0 if fused else self.substeps):
                _lib.call("altar_da_trajectory_propose", p(ws["phi"]), C, P, p(L), seed, stage, k, j,
                          p(ws["xi"]), p(ws["cand"]), s)
                ws["mask"].zero_()
                model.verify(candidate, ws["mask"])
                _lib.call("altar_mh_restore", p(ws["cand"]), p(ws["phi"]), p(ws["mask"]), C, P, s)
                model.priorLikelihood(candidate)
                sur.data_llk(ws["cand"], ws["cdata_s"], ws["z"])
                _lib.call("altar_da_trajectory_step", p(ws["phi"]), p(ws["prior_phi"]), p(ws["post_phi"]),
                          p(ws["cand"]), p(ws["cprior"]), p(ws["cdata_s"]), p(ws["cpost_s"]), p(ws["mask"]), beta,
                          C, P, seed, stage, k, j, p(ws["index"]), p(sub_counters), s)
            _lib.call("altar_da_trajectory_list", p(ws["index"]), C, p(ws["count"]), s)
            n = int(ws["count"].item())
            if n > 0:
                _lib.call("altar_da_pack", p(ws["phi"]), p(ws["index"]), n, P, p(ws["packed"]), s)
                packed = ws["packed"][:n]
                model.source.data_llk(model.restrict(packed), ws["cdata"][:n], model.precision)
                _lib.call("altar_da_accept", p(step.theta), p(step.prior), p(step.data), p(step.posterior),
                          p(ws["post_s"]), p(ws["phi"]), p(ws["prior_phi"]), p(ws["cdata"]), p(ws["post_phi"]),
                          p(ws["index"]), n, beta, P, seed, stage, k, None, p(counters), s)
            self.exact_evals += n
            moved_total += n
            d.notify(d.chainAdvanceFinish, annealer)
        acc, _, unl2, _ = (int(v) for v in counters.cpu().tolist())
        for i, v in enumerate(sub_counters.cpu().tolist()):
            self.substep_counts[i] += int(v)
        return acc, C * self.steps - moved_total, unl2
