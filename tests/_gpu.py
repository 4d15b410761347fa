"""Shared helpers for GPU tests: the engine module, the oracle, seeded inputs."""
import numpy as np

import omr_ntt as E
import oracle as O


def inputs(field, n, seed):
    return O.uniform(field.id, n, seed)


def gpu_ntt(field, x, variant=E.ARK_RADIX2, s=0):
    t = E.to_device(x, field)
    E.ntt_device(t, t, field, variant, s=s)
    return E.from_device(t, field)


def oracle_ntt(field, x):
    return O.ntt(field.id, x)
