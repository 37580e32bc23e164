"""Binning analysis for Monte Carlo time series (test infrastructure).

Textbook log-binning: the standard error of the mean is estimated at
successive bin sizes 2^k; the reported error is the largest estimate among the
levels that still have >= 64 bins (it plateaus once bins exceed the
autocorrelation time)."""

import numpy as np


def binning_error(x, min_bins=64):
    x = np.asarray(x, float)
    best = 0.0
    while len(x) >= min_bins:
        best = max(best, x.std(ddof=1) / np.sqrt(len(x)))
        n = len(x) // 2 * 2
        x = 0.5 * (x[:n:2] + x[1:n:2])
    return best


def mean_err(x):
    return float(np.mean(x)), binning_error(x)


def ratio_mean_err(num, den):
    """mean/error of <num>/<den> by jackknife over 64 blocks"""
    num = np.asarray(num, float)
    den = np.asarray(den, float)
    nb = 64
    n = len(num) // nb * nb
    nb_num = num[:n].reshape(nb, -1).mean(axis=1)
    nb_den = den[:n].reshape(nb, -1).mean(axis=1)
    tot_num, tot_den = nb_num.sum(), nb_den.sum()
    jk = (tot_num - nb_num) / (tot_den - nb_den)
    m = tot_num / tot_den
    err = np.sqrt((nb - 1) * np.mean((jk - jk.mean()) ** 2))
    return float(m), float(err)


def sigmas(a, ea, b, eb=0.0):
    """distance in combined standard errors"""
    return abs(a - b) / np.sqrt(ea * ea + eb * eb)


# observables of one sweep from the raw collector record (worker commit,
# energy.h:76-80, susceptibility.h:200-256, stiffness.h:123-128)
def observables(rec, beta, V, dim):
    rec = np.asarray(rec)
    E = rec[:, 2]
    return {
        "Energy Density": E / V,
        "Energy^2": E * E - rec[:, 0] / beta ** 2,
        "Energy": E,
        "Susceptibility": beta * rec[:, 9] / V,
        "Staggered Susceptibility": beta * rec[:, 16] / V,
        "Magnetization^2": rec[:, 5],
        "Staggered Magnetization^2": rec[:, 12],
        "Staggered Structure Factor": rec[:, 12] / V,
        "Number of Clusters": rec[:, 1],
        "Stiffness": rec[:, 17] / (beta * dim),
    }


# SSE (sse.C): the same record with the tau-integrated sums in units of the
# string; the susceptibilities are beta (n S + S0) / (n + 1) / V with n the
# operator count (susceptibility.h:213-251)
def observables_sse(rec, beta, V, dim):
    rec = np.asarray(rec)
    out = observables(rec, beta, V, dim)
    n = rec[:, 0]
    out["Susceptibility"] = beta * (n * rec[:, 9] + rec[:, 5]) / (n + 1) / V
    out["Staggered Susceptibility"] = beta * (n * rec[:, 16] + rec[:, 12]) / (n + 1) / V
    return out
