"""Mass-action reaction networks and the seeded synthetic generators.

A :class:`ReactionNetwork` is the model description this backend compiles: it
plays the role of an SBML model with mass-action kinetic laws imported by
AMICI's ``SbmlImporter`` (python/sdist/amici/importers/sbml/__init__.py:541-709,
``xdot = S w`` at :620-627).  Every reaction has at most two reactants and a
rate constant that is a free parameter; free parameters are log10-scaled
(``ParameterScaling.log10``, include/amici/defines.h:124).  Initial
concentrations are constants or fixed parameters (AMICI's
``fixed_parameters``); observables are linear combinations of species with
Gaussian or log-normal noise (python/sdist/amici/importers/utils.py:434-440).

Generators (all seeded with ``numpy.random.default_rng``):

* :func:`robertson` — Robertson kinetics A->B, 2B->B+C, B+C->A+C.
* :func:`synthetic_network` — random mass-action networks (40- and
  200-species presets in :func:`network_n40` / :func:`network_n200`).
"""
from __future__ import annotations

import copy
import dataclasses
import functools
import math
from typing import Sequence

import numpy as np

# Noise models of the observables (the continuous ones of AMICI's
# noise_distribution, python/sdist/amici/importers/utils.py:434-461): the
# distribution of the measurement m around the observable y with scale sigma,
# on the linear, natural-log or log10 scale.
NOISE_NORMAL = 0
NOISE_LOGNORMAL = 1
NOISE_LOG10NORMAL = 2
NOISE_LAPLACE = 3
NOISE_LOGLAPLACE = 4
NOISE_LOG10LAPLACE = 5
NOISE_NAMES = {"normal": NOISE_NORMAL, "lin-normal": NOISE_NORMAL, "log-normal": NOISE_LOGNORMAL,
               "lognormal": NOISE_LOGNORMAL, "log10-normal": NOISE_LOG10NORMAL,
               "laplace": NOISE_LAPLACE, "lin-laplace": NOISE_LAPLACE,
               "log-laplace": NOISE_LOGLAPLACE, "log10-laplace": NOISE_LOG10LAPLACE}


def noise_scale(noise: int) -> int:
    """Observable scaling of a noise model: 0 linear, 1 natural log, 2 log10
    (ObservableScaling, used by the residuals)."""
    return {NOISE_NORMAL: 0, NOISE_LAPLACE: 0, NOISE_LOGNORMAL: 1, NOISE_LOGLAPLACE: 1,
            NOISE_LOG10NORMAL: 2, NOISE_LOG10LAPLACE: 2}[int(noise)]


def noise_is_laplace(noise: int) -> bool:
    return int(noise) in (NOISE_LAPLACE, NOISE_LOGLAPLACE, NOISE_LOG10LAPLACE)


@dataclasses.dataclass
class Reaction:
    reactants: tuple  # species indices, multiplicity by repetition (<= 2)
    products: tuple
    rate: int  # index of the free parameter holding the rate constant


@dataclasses.dataclass
class ReactionNetwork:
    name: str
    species: list
    reactions: list
    parameter_names: list
    theta_nominal: np.ndarray  # log10 rate constants
    fixed_parameter_names: list
    k_nominal: np.ndarray
    x0_const: np.ndarray  # [nx]
    x0_fixed: np.ndarray  # [nx] int, -1 or index into fixed parameters
    observable_matrix: np.ndarray  # [ny, nx]
    noise: int = NOISE_NORMAL

    @property
    def nx(self) -> int:
        return len(self.species)

    @property
    def nr(self) -> int:
        return len(self.reactions)

    @property
    def np(self) -> int:
        return len(self.parameter_names)

    @property
    def nk(self) -> int:
        return len(self.fixed_parameter_names)

    @property
    def ny(self) -> int:
        return int(self.observable_matrix.shape[0])

    def stoichiometry(self) -> np.ndarray:
        s = np.zeros((self.nx, self.nr))
        for r, rx in enumerate(self.reactions):
            for a in rx.reactants:
                s[a, r] -= 1.0
            for b in rx.products:
                s[b, r] += 1.0
        return s

    def validate(self) -> None:
        for r, rx in enumerate(self.reactions):
            if len(rx.reactants) > 2:
                raise ValueError(f"reaction {r}: more than two reactants")
            for i in tuple(rx.reactants) + tuple(rx.products):
                if not 0 <= i < self.nx:
                    raise ValueError(f"reaction {r}: species index {i} out of range")
            if not 0 <= rx.rate < self.np:
                raise ValueError(f"reaction {r}: parameter index out of range")
        if self.observable_matrix.shape != (self.ny, self.nx):
            raise ValueError("observable matrix shape")
        if len(self.x0_const) != self.nx or len(self.x0_fixed) != self.nx:
            raise ValueError("initial condition arrays must have nx entries")

    # flattened tables shared with the oracle harness (oracle/amx_oracle.h)
    def flat_tables(self) -> dict:
        rptr, ridx, pptr, pidx = [0], [], [0], []
        for rx in self.reactions:
            ridx.extend(rx.reactants)
            rptr.append(len(ridx))
            pidx.extend(rx.products)
            pptr.append(len(pidx))
        return dict(
            rptr=np.asarray(rptr, np.int32),
            ridx=np.asarray(ridx if ridx else [0], np.int32),
            pptr=np.asarray(pptr, np.int32),
            pidx=np.asarray(pidx if pidx else [0], np.int32),
            kidx=np.asarray([rx.rate for rx in self.reactions], np.int32),
            x0c=np.asarray(self.x0_const, np.float64),
            x0k=np.asarray(self.x0_fixed, np.int32),
            C=np.ascontiguousarray(self.observable_matrix, np.float64),
        )


def robertson() -> ReactionNetwork:
    """Robertson kinetics as three mass-action reactions.

    A -> B (k1 = 0.04), 2B -> B + C (k2 = 3e7), B + C -> A + C (k3 = 1e4),
    x0 = (1, 0, 0); observables y = (A, C).  (AMICI's own test model
    python/sdist/amici/testing/models.py:90-110 orders the constants
    (0.04, 1e4, 3e7); here they follow the reaction order above.)
    """
    rx = [Reaction((0,), (1,), 0), Reaction((1, 1), (1, 2), 1), Reaction((1, 2), (0, 2), 2)]
    c = np.zeros((2, 3))
    c[0, 0] = 1.0
    c[1, 2] = 1.0
    return ReactionNetwork(
        name="robertson",
        species=["A", "B", "C"],
        reactions=rx,
        parameter_names=["k1", "k2", "k3"],
        theta_nominal=np.log10([0.04, 3e7, 1e4]),
        fixed_parameter_names=[],
        k_nominal=np.zeros(0),
        x0_const=np.array([1.0, 0.0, 0.0]),
        x0_fixed=np.full(3, -1, np.int32),
        observable_matrix=c,
        noise=NOISE_NORMAL,
    )


def synthetic_network(
    n_species: int,
    n_reactions: int,
    seed: int,
    frac_unimolecular: float = 0.6,
    log10_k_range: tuple = (-3.0, 4.0),
    n_inputs: int = 8,
    n_observables: int = 30,
    terms_per_observable: int = 3,
    x0_base: float = 1.0,
    noise: int = NOISE_LOGNORMAL,
    name: str | None = None,
    t_end: float = 100.0,
    min_observed: float = 1e-3,
) -> ReactionNetwork:
    """Seeded random mass-action network.

    * unimolecular reactions A -> B with B != A;
    * bimolecular reactions A + B -> C + D (A == B allowed, i.e. 2A), never
      with the product multiset equal to the reactant multiset.
    Every reaction keeps the molecule count, so trajectories stay bounded and
    non-negative.  Rate constants are log-uniform in ``10**log10_k_range``.
    ``n_inputs`` distinct species take their initial concentration from a
    fixed parameter; the others start at ``x0_base``.  Observables are sums of
    ``terms_per_observable`` distinct species with weights U(0.5, 2), drawn
    from the species whose concentration stays >= ``min_observed`` on
    [0, t_end] at the nominal rates for inputs 0.1, 1 and 10 (determined once
    with scipy's stiff integrator), so that log-normal observables stay away
    from the absolute-tolerance floor.
    """
    rng = np.random.default_rng(seed)
    n_uni = int(round(frac_unimolecular * n_reactions))
    reactions = []
    for r in range(n_reactions):
        if r < n_uni:
            a = int(rng.integers(n_species))
            b = int(rng.integers(n_species - 1))
            b += int(b >= a)
            reactions.append(Reaction((a,), (b,), r))
        else:
            while True:
                re = tuple(sorted(int(v) for v in rng.integers(n_species, size=2)))
                pr = tuple(sorted(int(v) for v in rng.integers(n_species, size=2)))
                if re != pr:
                    break
            reactions.append(Reaction(re, pr, r))
    theta = rng.uniform(log10_k_range[0], log10_k_range[1], size=n_reactions)
    inputs = np.sort(rng.choice(n_species, size=n_inputs, replace=False))
    x0c = np.full(n_species, x0_base)
    x0k = np.full(n_species, -1, np.int32)
    for j, s in enumerate(inputs):
        x0k[s] = j
        x0c[s] = 0.0
    probe = ReactionNetwork(
        name="probe", species=[f"X{i}" for i in range(n_species)], reactions=reactions,
        parameter_names=[f"k{r}" for r in range(n_reactions)], theta_nominal=theta,
        fixed_parameter_names=[f"x0_X{int(s)}" for s in inputs], k_nominal=np.ones(n_inputs),
        x0_const=x0c, x0_fixed=x0k, observable_matrix=np.zeros((0, n_species)), noise=noise)
    floor = _species_floor(probe, t_end)
    cand = np.where(floor >= min_observed)[0]
    if len(cand) < terms_per_observable + 2:
        cand = np.argsort(floor)[::-1][: max(terms_per_observable + 2, len(cand))]
    c = np.zeros((n_observables, n_species))
    for k in range(n_observables):
        sp = rng.choice(cand, size=terms_per_observable, replace=False)
        c[k, sp] = rng.uniform(0.5, 2.0, size=terms_per_observable)
    net = ReactionNetwork(
        name=name or f"massaction_n{n_species}_r{n_reactions}_s{seed}",
        species=[f"X{i}" for i in range(n_species)],
        reactions=reactions,
        parameter_names=[f"k{r}" for r in range(n_reactions)],
        theta_nominal=theta,
        fixed_parameter_names=[f"x0_X{int(s)}" for s in inputs],
        k_nominal=np.ones(n_inputs),
        x0_const=x0c,
        x0_fixed=x0k,
        observable_matrix=c,
        noise=noise,
    )
    net.validate()
    return net


def _species_floor(net: ReactionNetwork, t_end: float) -> np.ndarray:
    """Minimum concentration of every species over [0, t_end] at the nominal
    rates, over input levels 0.1, 1 and 10 (scipy BDF, rtol 1e-6)."""
    from scipy.integrate import solve_ivp

    S = net.stoichiometry()
    p = 10.0 ** np.asarray(net.theta_nominal)
    react = [list(rx.reactants) for rx in net.reactions]
    kidx = np.array([rx.rate for rx in net.reactions])

    def f(_t, x):
        w = p[kidx].copy()
        for r, re in enumerate(react):
            for a in re:
                w[r] *= x[a]
        return S @ w

    def jac(_t, x):
        J = np.zeros((net.nx, net.nx))
        for r, re in enumerate(react):
            for i, a in enumerate(re):
                d = p[kidx[r]]
                for j, b in enumerate(re):
                    if j != i:
                        d *= x[b]
                J[:, a] += S[:, r] * d
        return J

    floor = np.full(net.nx, np.inf)
    for level in (0.1, 1.0, 10.0):
        x0 = np.where(net.x0_fixed >= 0, level, net.x0_const)
        sol = solve_ivp(f, (0.0, t_end), x0, method="BDF", jac=jac, rtol=1e-6, atol=1e-12,
                        t_eval=np.linspace(0.0, t_end, 201))
        floor = np.minimum(floor, sol.y.min(axis=1))
    return floor


@functools.lru_cache(maxsize=None)
def _n40_cached() -> ReactionNetwork:
    return synthetic_network(40, 120, seed=3, log10_k_range=(-3.0, 4.0), n_inputs=8,
                             n_observables=30, name="n40")


def network_n40() -> ReactionNetwork:
    """40 species, 120 reactions, k log-uniform 1e-3..1e4, 8 inputs, 30 observables (seed 3)."""
    return copy.deepcopy(_n40_cached())


def network_n200() -> ReactionNetwork:
    """200 species, 600 reactions, k log-uniform 1e-4..1e5, 20 inputs, 60 observables (seed 4)."""
    return synthetic_network(200, 600, seed=4, log10_k_range=(-4.0, 5.0), n_inputs=20,
                             n_observables=60, name="n200", t_end=500.0)


def open_network(base: ReactionNetwork, seed: int, log10_kdeg_range: tuple = (-2.0, 0.0),
                 n_synthesis: int = 5, log10_ksyn_range: tuple = (-1.0, 1.0),
                 name: str | None = None) -> ReactionNetwork:
    """``base`` made an open system: every species degrades (X -> 0, its own
    rate constant, log-uniform in ``10**log10_kdeg_range``) and
    ``n_synthesis`` species are produced at constant rates (0 -> X).  The
    stoichiometric matrix then has full row rank (no conserved moieties), so
    the steady state is isolated and its Jacobian is non-singular — the
    setting of pre-equilibration with Newton's method and implicit-function
    sensitivities (src/steadystateproblem.cpp, src/backwardproblem.cpp:394-460).
    The input species, observables and noise model of ``base`` are kept."""
    rng = np.random.default_rng(seed)
    rx = list(copy.deepcopy(base.reactions))
    names = list(base.parameter_names)
    theta = list(base.theta_nominal)
    for i in range(base.nx):
        rx.append(Reaction((i,), (), len(names)))
        names.append(f"kdeg_{base.species[i]}")
        theta.append(rng.uniform(*log10_kdeg_range))
    for i in np.sort(rng.choice(base.nx, size=n_synthesis, replace=False)):
        rx.append(Reaction((), (int(i),), len(names)))
        names.append(f"ksyn_{base.species[int(i)]}")
        theta.append(rng.uniform(*log10_ksyn_range))
    net = dataclasses.replace(copy.deepcopy(base), name=name or base.name + "_open", reactions=rx,
                              parameter_names=names, theta_nominal=np.asarray(theta))
    net.validate()
    return net


def network_n40_open() -> ReactionNetwork:
    """The 40-species network with degradation of every species and synthesis
    of five (175 reactions, 165 rate constants; seed 5)."""
    return open_network(network_n40(), seed=5, name="n40open")


def network_n200_open() -> ReactionNetwork:
    """The 200-species network with degradation of every species and
    synthesis of five (805 reactions, 805 rate constants; seed 6)."""
    return open_network(network_n200(), seed=6, name="n200open")


def linear_chain(n: int = 3, seed: int = 11) -> ReactionNetwork:
    """Closed linear chain X0 <-> X1 <-> ... (first-order reactions only).

    xdot = A x with A built from the rate constants, so states and every
    sensitivity have the closed form expm(A t) x0 — used to pin both the oracle
    and the kernels.
    """
    rng = np.random.default_rng(seed)
    rx = []
    for i in range(n - 1):
        rx.append(Reaction((i,), (i + 1,), len(rx)))
        rx.append(Reaction((i + 1,), (i,), len(rx)))
    theta = rng.uniform(-1.0, 0.5, size=len(rx))
    c = np.eye(n)[: min(2, n)]
    x0 = np.zeros(n)
    x0[0] = 1.0
    return ReactionNetwork(
        name=f"linear{n}",
        species=[f"X{i}" for i in range(n)],
        reactions=rx,
        parameter_names=[f"k{r}" for r in range(len(rx))],
        theta_nominal=theta,
        fixed_parameter_names=[],
        k_nominal=np.zeros(0),
        x0_const=x0,
        x0_fixed=np.full(n, -1, np.int32),
        observable_matrix=c,
        noise=NOISE_NORMAL,
    )


# ---------------------------------------------------------------------------
# condition sets of the benchmark configurations

def with_noise(net: ReactionNetwork, noise: int) -> ReactionNetwork:
    """The same network with another noise model (its own model module)."""
    if int(noise) == int(net.noise):
        return net
    return dataclasses.replace(net, name=f"{net.name}_noise{int(noise)}", noise=int(noise))


def robertson_times() -> np.ndarray:
    """40 log-spaced output times 1e-4 .. 4e5."""
    return np.logspace(-4.0, math.log10(4e5), 40)


def robertson_parameter_sets(n: int = 4096, seed: int = 2) -> np.ndarray:
    """log10 k_i ~ Normal(log10 nominal, 0.5), [n, 3]."""
    rng = np.random.default_rng(seed)
    nom = robertson().theta_nominal
    return nom[None, :] + 0.5 * rng.standard_normal((n, 3))


def input_conditions(net: ReactionNetwork, n: int, seed: int) -> np.ndarray:
    """Input initial concentrations log-uniform 0.1 .. 10, [n, nk]."""
    rng = np.random.default_rng(seed)
    return 10.0 ** rng.uniform(-1.0, 1.0, size=(n, net.nk))


def add_noise(y: np.ndarray, noise: int, sigma: float, seed: int) -> np.ndarray:
    """Synthetic measurements drawn from the noise model around y: Gaussian
    or Laplace on the linear, natural-log or log10 scale."""
    rng = np.random.default_rng(seed)
    e = rng.laplace(size=y.shape) if noise_is_laplace(noise) else rng.standard_normal(y.shape)
    sc = noise_scale(noise)
    if sc == 1:
        return y * np.exp(sigma * e)
    if sc == 2:
        return y * 10.0 ** (sigma * e)
    return y + sigma * e


def condition_set(name: str) -> dict:
    """Inputs of a benchmark configuration except the measurements.

    Returns network, theta [B, np], kfix [B, nk], ts [nt], sigma (scalar) and
    the seed used for the measurement noise.
    """
    if name in ("R", "RB"):
        net = robertson()
        nb = 1 if name == "R" else 4096
        theta = (net.theta_nominal[None, :].repeat(1, 0) if name == "R"
                 else robertson_parameter_sets(nb, seed=2))
        return dict(network=net, theta=theta, kfix=np.zeros((nb, 0)), ts=robertson_times(),
                    sigma=0.01, noise_seed=2)
    if name == "N40":
        net = network_n40()
        nb = 1024
        return dict(network=net, theta=np.repeat(net.theta_nominal[None, :], nb, 0),
                    kfix=input_conditions(net, nb, seed=3), ts=np.linspace(0.0, 100.0, 25),
                    sigma=0.1, noise_seed=3)
    if name == "N40P":
        # pre-equilibrated: steady state at input level 1, then the inputs are
        # reset to the condition's levels at t = 0
        net = network_n40_open()
        nb = 1024
        return dict(network=net, theta=np.repeat(net.theta_nominal[None, :], nb, 0),
                    kfix=input_conditions(net, nb, seed=5), kfix_preeq=np.ones((nb, net.nk)),
                    ts=np.linspace(0.0, 100.0, 25), sigma=0.1, noise_seed=5)
    if name == "N200P":
        # pre-equilibrated as N40P: steady state at input level 1, inputs reset
        net = network_n200_open()
        nb = 256
        return dict(network=net, theta=np.repeat(net.theta_nominal[None, :], nb, 0),
                    kfix=input_conditions(net, nb, seed=6), kfix_preeq=np.ones((nb, net.nk)),
                    ts=np.linspace(0.0, 500.0, 40), sigma=0.1, noise_seed=6)
    if name == "N200":
        net = network_n200()
        nb = 256
        return dict(network=net, theta=np.repeat(net.theta_nominal[None, :], nb, 0),
                    kfix=input_conditions(net, nb, seed=4), ts=np.linspace(0.0, 500.0, 40),
                    sigma=0.1, noise_seed=4)
    raise KeyError(name)


def mix_groups(total: int = 2048, seed: int = 5) -> list:
    """The mixed workload: ``total`` conditions drawn 50 / 35 / 15 % from the
    Robertson, 40-species and 200-species generators.  The two reaction
    networks are the open variants, pre-equilibrated to steady state at input
    level 1 (their closed forms have conserved moieties and no isolated steady
    state); Robertson has no fixed parameters, so its conditions cannot be
    pre-equilibrated and are simulated from the initial state.  Returns one
    condition set per model type (as :func:`condition_set`)."""
    n_r = total // 2
    n_40 = int(round(0.35 * total))
    n_200 = total - n_r - n_40
    net_r = robertson()
    n40, n200 = network_n40_open(), network_n200_open()
    return [
        dict(network=net_r, theta=robertson_parameter_sets(n_r, seed=seed), kfix=np.zeros((n_r, 0)),
             ts=robertson_times(), sigma=0.01, noise_seed=seed),
        dict(network=n40, theta=np.repeat(n40.theta_nominal[None, :], n_40, 0),
             kfix=input_conditions(n40, n_40, seed=seed), kfix_preeq=np.ones((n_40, n40.nk)),
             ts=np.linspace(0.0, 100.0, 25), sigma=0.1, noise_seed=seed),
        dict(network=n200, theta=np.repeat(n200.theta_nominal[None, :], n_200, 0),
             kfix=input_conditions(n200, n_200, seed=seed), kfix_preeq=np.ones((n_200, n200.nk)),
             ts=np.linspace(0.0, 500.0, 40), sigma=0.1, noise_seed=seed),
    ]


def pick(values: Sequence, idx: Sequence[int]) -> list:
    return [values[i] for i in idx]
