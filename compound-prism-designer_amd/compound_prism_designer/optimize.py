"""Generation-level driver for the designer's optimisation loop.

The reference optimises with pymoo, one individual per `_evaluate` call
(designer.py:115-300: CompoundPrismSpectrometerProblem builds a Spectrometer
from each parameter vector and calls gpu_fitness on it).  Here a whole
population is evaluated per call: its parameter vectors become `cpd_params`
records, the designs are built on the GPU (detectors positioned, or placed
from the exit face) and evaluated by one of three estimators:

- "plan": the fixed-count fitness in one captured generation
  (cpd_plan_run_params: construction and evaluation replay as HIP graphs);
- "server": the same fitness through a continuous-batching server
  (cpd_server_*), so several problems or callers share the GPU and their
  generations overlap;
- "gpu_fitness": the designer's own estimator, designer.py:276
  (gpu_fitness with max_n 128 and max_eval 8192, cpd_gpu_fitness_batch).

pymoo is not part of this build, so `evolve` is a small elitist
multi-objective loop of its own (non-dominated sorting with crowding
distance, simulated binary crossover, polynomial mutation, the designer's
feasibility constraint); any other optimiser can call
`SpectrometerProblem.evaluate` on its populations instead.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import Context, Glass, Plan, Server, _check, build_designs, gpu_fitness_batch
from . import _native as N


def sellmeier_glass(name: str, b: Sequence[float], c: Sequence[float], wavelength_range=(0.5, 0.82),
                    npts: int = 100) -> Glass:
    """A Sellmeier1 glass fitted to the Horner polynomial the kernels evaluate
    (cpd_fit_sellmeier1; AbstractGlass.into_glass, glasscat.py:114-127)."""
    bb = (ctypes.c_double * 3)(*b)
    cc = (ctypes.c_double * 3)(*c)
    out = (ctypes.c_double * 6)()
    _check(N.lib.cpd_fit_sellmeier1(bb, cc, float(wavelength_range[0]), float(wavelength_range[1]), npts, out,
                                    None))
    return Glass(name, list(out))


def fit_sellmeier_catalog(ctx: Context, b, c, wavelength_range=(0.5, 0.82), npts: int = 100,
                          names: Optional[Sequence[str]] = None):
    """Fits n Sellmeier1 glasses (b, c: n x 3) on the GPU in one call, one glass
    per thread (cpd_fit_sellmeier1_batch; glasscat.py:114-127 per glass).
    Returns (glasses, chisq); a glass whose fit fails (a pole in range) raises
    ValueError naming it."""
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
    c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1, 3)
    if b.shape != c.shape:
        raise ValueError("b and c must have the same number of glasses")
    n = b.shape[0]
    names = list(names) if names is not None else [f"G{i:03d}" for i in range(n)]
    coeffs = np.empty((n, 6))
    chisq = np.empty(n)
    status = (ctypes.c_int * max(n, 1))()
    dp = ctypes.POINTER(ctypes.c_double)
    _check(N.lib.cpd_fit_sellmeier1_batch(ctx.handle, b.ctypes.data_as(dp), c.ctypes.data_as(dp), n,
                                          float(wavelength_range[0]), float(wavelength_range[1]), npts,
                                          coeffs.ctypes.data_as(dp), chisq.ctypes.data_as(dp), status))
    bad = [names[i] for i in range(n) if status[i] != N.OK]
    if bad:
        raise ValueError(f"Sellmeier fit failed (pole in range) for {bad[:8]}")
    return [Glass(names[i], list(coeffs[i])) for i in range(n)], chisq


def synthetic_catalog(n: int, seed: int = 0, wavelength_range=(0.5, 0.82)) -> List[Glass]:
    """n Sellmeier1 glasses with B ~ U[0.3, 1.5], C ~ U[0.001, 0.2] um^2 (the
    benchmark generator's ranges), standing in for a glass catalogue file."""
    rng = np.random.default_rng(seed)
    return [sellmeier_glass(f"G{i:03d}", rng.uniform(0.3, 1.5, 3), rng.uniform(0.001, 0.2, 3), wavelength_range)
            for i in range(n)]


_PARAMS_DTYPE = None


def _params_dtype():
    """numpy view type of cpd_params records (built once: it is slow to derive)."""
    global _PARAMS_DTYPE
    if _PARAMS_DTYPE is None:
        _PARAMS_DTYPE = np.ctypeslib.as_array((N.Params * 1)()).dtype
    return _PARAMS_DTYPE


@dataclass
class SpectrometerProblem:
    """The designer's problem (designer.py:115-201), Gaussian beam.  Variables
    are normalised to [0, 1]: glass indices (nglass), face angles (nglass + 1,
    mapped to [-pi/2, pi/2]), tanh of the separation lengths (nglass; length =
    artanh(t) max_height / 4, designer.py:217), curvature ([1e-5, 1]), height
    ([1e-4, 1] max_height), normalised beam height ([0, 1]) and detector angle
    ([-pi, pi]).  With acceptance None (the designer's
    auto_position_detector_acceptance unset) two more variables place the
    detector from the exit face: distance ([0, 40] max_height) and angle
    ([-pi/2, pi/2]) (designer.py:131-132, 236-243); otherwise the detector is
    positioned for that acceptance.

    estimator: "plan", "server" or "gpu_fitness" (module docstring).  The
    designer draws one random seed per individual for gpu_fitness; here one
    seed per generation serves all its individuals, since a seed fixes the
    beam table that every design of the launch reads.  A design that does not
    converge gets the designer's RayTraceError objectives."""

    glasses: Sequence[Glass]
    nglass: int
    max_height: float = 25.0
    prism_width: float = 7.0
    ar_coated: bool = True
    beam_width: float = 3.2
    wavelength_range: tuple = (0.5, 0.82)
    bin_count: int = 32
    bin_size: float = 0.8
    linear_slope: float = 1.0
    linear_intercept: float = 0.1
    detector_length: float = 32.0
    max_incident_angle: float = math.radians(45.0)
    acceptance: Optional[float] = 1.0
    n_wavelengths: int = 64
    n_rays: int = 16384
    optimize_size: bool = True
    optimize_deviation: bool = True
    ctx: Optional[Context] = None
    estimator: str = "plan"
    gpu_fitness_max_n: int = 128
    gpu_fitness_max_eval: int = 16_384 // 2
    seed: int = 0
    server: Optional[Server] = None
    _plan: Optional[Plan] = field(default=None, init=False, repr=False)
    _own_server: bool = field(default=False, init=False, repr=False)
    _rng: Optional[np.random.Generator] = field(default=None, init=False, repr=False)

    def __post_init__(self):
        if self.estimator not in ("plan", "server", "gpu_fitness"):
            raise ValueError(f"unknown estimator {self.estimator!r}")
        if self.acceptance is not None and not 0.0 < self.acceptance <= 1.0:
            raise ValueError("acceptance must be in (0, 1] or None")
        self._rng = np.random.default_rng(self.seed)

    @property
    def explicit_position(self) -> bool:
        return self.acceptance is None

    @property
    def build_acceptance(self) -> float:
        """The construction's acceptance argument (cpd_build_designs)."""
        return N.POSITION_FROM_EXIT if self.acceptance is None else float(self.acceptance)

    @property
    def n_var(self) -> int:
        return 3 * self.nglass + 1 + 4 + (2 if self.explicit_position else 0)

    @property
    def n_obj(self) -> int:
        return 1 + int(self.optimize_size) + int(self.optimize_deviation)

    def params(self, X: np.ndarray):
        """The cpd_params records of a population X [pop, n_var] as a ctypes
        array (create_spectrometer, designer.py:203-265, with the detector left
        to positioning), filled through a numpy view of the records."""
        X = np.clip(np.atleast_2d(np.asarray(X, np.float64)), 0.0, 1.0)
        n, g, P = len(X), len(self.glasses), self.nglass
        arr = (N.Params * n)()
        a = np.frombuffer(arr, dtype=_params_dtype())
        coeffs = np.array([gl.glass for gl in self.glasses], np.float64)
        idx = np.minimum((X[:, :P] * g).astype(np.int64), g - 1)
        a["n_glass"] = P
        a["ar_coated"] = int(self.ar_coated)
        a["glass"][:, :P, :] = coeffs[idx]
        a["angles"][:, :P + 1] = -math.pi / 2 + math.pi * X[:, P:2 * P + 1]
        a["lengths"][:, :P] = np.arctanh(np.minimum(X[:, 2 * P + 1:3 * P + 1], 1.0 - 1e-12)) * self.max_height / 4
        a["curvature"] = 1e-5 + (1.0 - 1e-5) * X[:, 3 * P + 1]
        a["height"] = self.max_height * (1e-4 + (1.0 - 1e-4) * X[:, 3 * P + 2])
        a["y_mean"] = a["height"] * X[:, 3 * P + 3]
        a["det_angle"] = -math.pi + 2 * math.pi * X[:, 3 * P + 4]
        a["width"] = self.prism_width
        a["bin_count"] = self.bin_count
        a["bin_size"] = self.bin_size
        a["linear_slope"] = self.linear_slope
        a["linear_intercept"] = self.linear_intercept
        a["det_length"] = self.detector_length
        a["max_incident_angle"] = self.max_incident_angle
        a["beam_width"] = self.beam_width
        a["wavelength_lo"], a["wavelength_hi"] = self.wavelength_range
        if self.explicit_position:
            a["det_position"][:, 0] = self.max_height * 40 * X[:, 3 * P + 5]
            a["det_position"][:, 1] = -math.pi / 2 + math.pi * X[:, 3 * P + 6]
            a["flipped"] = 0
        return arr

    def _plan_for(self, params: List["N.Params"]) -> Plan:
        n = len(params)
        if self._plan is not None and self._plan.n == n:
            return self._plan
        if self._plan is not None:
            self._plan.close()
        # every slot has the same shape (sampling, bin count): any buildable
        # individual is a valid stand-in record for all of them; random
        # parameter vectors mostly fail (total internal reflection), so draw
        # more when the population has none
        designs, st = build_designs(params, self.build_acceptance, ctx=self.ctx)
        rng = np.random.default_rng(0)
        for _ in range(64):
            if (st == 0).any():
                break
            designs, st = build_designs(self.params(rng.random((256, self.n_var))), self.build_acceptance,
                                        ctx=self.ctx)
        ok = np.flatnonzero(st == 0)
        if ok.size == 0:
            raise RuntimeError("no buildable design found for the plan's stand-in record")
        shape = [designs[int(ok[0])]] * n
        # the default schedule: tuning on n copies of one stand-in says nothing
        # about the populations to come
        self._plan = Plan(shape, (self.n_wavelengths, self.n_rays), ctx=self.ctx,
                          schedule=N.Schedule(0, 16, N.ORDER_GIVEN, 0))
        return self._plan

    def _built(self, params):
        """Designs built on the GPU; (designs, status, indices of the built ones)."""
        designs, st = build_designs(params, self.build_acceptance, ctx=self.ctx)
        return designs, st, np.flatnonzero(st == 0)

    def _server_fitness(self, params):
        designs, st, ok = self._built(params)
        fit = np.full((len(params), 3), np.nan)
        if ok.size:
            if self.server is None:
                self.server = Server(self.ctx)
                self._own_server = True
            rid = self.server.submit([designs[i] for i in ok], (self.n_wavelengths, self.n_rays))
            res, rst, _ = self.server.wait(rid)
            fit[ok] = res
            st = st.copy()
            st[ok] = rst
        return fit, st

    def _gpu_fitness(self, params):
        designs, st, ok = self._built(params)
        fit = np.full((len(params), 3), np.nan)
        if ok.size:
            seeds = self._rng.random(1)
            res, rst = gpu_fitness_batch([designs[i] for i in ok], seeds, max_n=self.gpu_fitness_max_n,
                                         max_eval=self.gpu_fitness_max_eval, ctx=self.ctx)
            fit[ok] = res
            st = st.copy()
            st[ok] = rst
        return fit, st

    def fitness(self, X: np.ndarray):
        """Raw fitness [pop, 3] (size, info, deviation) and per-design status of
        a population, by the problem's estimator."""
        params = self.params(X)
        if self.estimator == "server":
            return self._server_fitness(params)
        if self.estimator == "gpu_fitness":
            return self._gpu_fitness(params)
        return self._plan_for(params).run_params(params, self.build_acceptance)

    def evaluate(self, X: np.ndarray):
        """Objectives F [pop, n_obj] (size, log2(nbin) - info, deviation, as
        designer.py:267-300 orders them), constraint G [pop] (0 feasible),
        the raw fitness [pop, 3] and the per-design status."""
        fit, st = self.fitness(X)
        max_size = self.max_height * 40 if self.optimize_size else np.inf
        max_info = math.log2(self.bin_count)
        size, info, dev = fit[:, 0], fit[:, 1], fit[:, 2]
        with np.errstate(invalid="ignore"):
            good = (st == 0) & np.isfinite(info) & (max_info - info > 0)
            feasible = good & (self.max_height / 2 < size) & (size < max_size)
        cols = []
        if self.optimize_size:
            cols.append(np.where(good, size, max_size * 10))
        cols.append(np.where(good, max_info - info, max_info))
        if self.optimize_deviation:
            cols.append(np.where(good, dev, 1.0))
        F = np.stack(cols, axis=1)
        G = np.where(feasible, 0.0, 1.0)
        return F, G, fit, st

    def close(self):
        if self._plan is not None:
            self._plan.close()
            self._plan = None
        if self._own_server and self.server is not None:
            self.server.close()
            self.server = None
            self._own_server = False


def nondominated_ranks(F: np.ndarray) -> np.ndarray:
    """Front index of each row of F (minimisation; 0 = non-dominated)."""
    n = len(F)
    le = (F[:, None, :] <= F[None, :, :]).all(-1)
    lt = (F[:, None, :] < F[None, :, :]).any(-1)
    dom = le & lt  # dom[i, j]: i dominates j
    count = dom.sum(0)
    rank = np.full(n, -1)
    front = np.flatnonzero(count == 0)
    r = 0
    while front.size:
        rank[front] = r
        count = count - dom[front].sum(0)
        count[rank >= 0] = -1
        front = np.flatnonzero(count == 0)
        r += 1
    return rank


def crowding(F: np.ndarray) -> np.ndarray:
    n, m = F.shape
    d = np.zeros(n)
    if n <= 2:
        return np.full(n, np.inf)
    for k in range(m):
        o = np.argsort(F[:, k], kind="stable")
        span = F[o[-1], k] - F[o[0], k]
        d[o[0]] = d[o[-1]] = np.inf
        if span > 0:
            d[o[1:-1]] += (F[o[2:], k] - F[o[:-2], k]) / span
    return d


def survive(F: np.ndarray, G: np.ndarray, k: int) -> np.ndarray:
    """Indices of k survivors: feasible before infeasible (each group ranked on
    its own, so an infeasible row never dominates a feasible one), then by
    front, then by crowding distance (wider first)."""
    rank = np.zeros(len(F), np.int64)
    crowd = np.zeros(len(F))
    for g in np.unique(G):
        grp = np.flatnonzero(G == g)
        rank[grp] = nondominated_ranks(F[grp])
        for r in np.unique(rank[grp]):
            idx = grp[rank[grp] == r]
            crowd[idx] = crowding(F[idx])
    order = np.lexsort((-crowd, rank, G))
    return order[:k]


def _sbx(rng, a, b, eta):
    u = rng.random(a.shape)
    beta = np.where(u <= 0.5, (2 * u) ** (1 / (eta + 1)), (1 / (2 * (1 - u))) ** (1 / (eta + 1)))
    c1 = 0.5 * ((1 + beta) * a + (1 - beta) * b)
    c2 = 0.5 * ((1 - beta) * a + (1 + beta) * b)
    swap = rng.random(a.shape) < 0.5
    return np.clip(np.where(swap, c2, c1), 0, 1)


def _mutate(rng, x, eta, prob):
    u = rng.random(x.shape)
    delta = np.where(u < 0.5, (2 * u) ** (1 / (eta + 1)) - 1, 1 - (2 * (1 - u)) ** (1 / (eta + 1)))
    m = rng.random(x.shape) < prob
    return np.clip(np.where(m, x + delta * np.where(u < 0.5, x, 1 - x), x), 0, 1)


def evolve(problem: SpectrometerProblem, pop_size: int = 64, generations: int = 10, seed: int = 0,
           crossover_eta: float = 4.0, mutation_eta: float = 2.0, mutation_prob: Optional[float] = None,
           X0: Optional[np.ndarray] = None):
    """Elitist (mu + mu) multi-objective loop; every generation of pop_size
    offspring is one evaluate call (one captured generation, one server
    request or one batched gpu_fitness call, by the problem's estimator).  The distribution indices and
    the per-variable mutation probability default to design_config.toml's
    optimiser section (4, 2 and 0.05).  Returns the final population (X, F, G)
    and a per-generation history."""
    rng = np.random.default_rng(seed)
    prob = mutation_prob if mutation_prob is not None else 0.05
    X = rng.random((pop_size, problem.n_var)) if X0 is None else np.asarray(X0, np.float64)
    t = time.perf_counter()
    F, G, _, st = problem.evaluate(X)
    history = [{"generation": 0, "eval_s": time.perf_counter() - t, "feasible": int((G == 0).sum()),
                "built": int((st == 0).sum()), "best": F[G == 0].min(0).tolist() if (G == 0).any() else None}]
    for gen in range(1, generations + 1):
        # binary tournaments on (feasibility, front)
        rank = nondominated_ranks(F) + 1000 * G
        a, b = rng.integers(0, pop_size, (2, pop_size))
        parents = np.where(rank[a] <= rank[b], a, b)
        mates = np.roll(parents, 1)
        Y = _mutate(rng, _sbx(rng, X[parents], X[mates], crossover_eta), mutation_eta, prob)
        t = time.perf_counter()
        FY, GY, _, stY = problem.evaluate(Y)
        eval_s = time.perf_counter() - t
        XX, FF, GG = np.vstack([X, Y]), np.vstack([F, FY]), np.concatenate([G, GY])
        keep = survive(FF, GG, pop_size)
        X, F, G = XX[keep], FF[keep], GG[keep]
        history.append({"generation": gen, "eval_s": eval_s, "feasible": int((G == 0).sum()),
                        "built": int((stY == 0).sum()),
                        "best": F[G == 0].min(0).tolist() if (G == 0).any() else None})
    return X, F, G, history
