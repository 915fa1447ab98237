"""Shared builders for the tests (synthetic specs are cached per process)."""
import functools

import numpy as np

from blade_amd import synthetic, system


@functools.lru_cache(maxsize=None)
def spec(name):
    if name == "config1":
        return synthetic.config1()
    if name == "msld_small":
        return synthetic.msld_small()
    if name == "config2":
        return synthetic.config2()
    if name == "config2_small":
        return synthetic.config2(seed=0xB1ADE0B0, n_water=1200, n_poly=400, box=36.0)
    if name == "config3":
        return synthetic.config3()
    if name == "config5":
        return synthetic.config5(0)
    if name == "two_site_small":
        return synthetic.two_site_small()
    raise KeyError(name)


def built(name):
    s = system.build(spec(name))
    s.prepare()
    return s


@functools.lru_cache(maxsize=None)
def relaxed(name):
    """the spec advanced through the GPU relaxation schedule (needs a GPU)"""
    return system.relax(spec(name))


def lambdas0(s, name):
    """lambda of the initial theta by the oracle's rule"""
    from oracle import oracle as orc
    t = s.topology().contents
    if t.n_blocks <= 1:
        return np.ones(1)
    return orc.theta_to_lambda(s.topology(), s.runparams(), np.array(spec(name).theta0, dtype=float))


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
