"""The benchmark workloads as generator settings.

config1: 1 Mbp i.i.d. genome, 100 bp reads at 30x, 0.5% substitutions, k=31
config2: 100 Mbp genome with 2% from 50 repeat families (500-5000 bp, 1%
         divergence), 150 bp reads at 50x (~33.3 M reads), 1% substitutions, k=31
config3: 20 Mbp (same repeat layer), 250 bp reads at 40x, 0.5% substitutions
         + 0.05% 1-bp indels, k=63
config4: 256 independent jobs: random genomes log-uniform in 50 kbp - 20 Mbp,
         read length uniform in {100, 150, 250}, coverage uniform in 20-60x,
         1% substitutions, k=31 (run as batches)
config5: config4-distribution jobs with k = 31 or 63 (80/20) arriving as a
         Poisson process (served)
All use the -T -L 3 thresholds.
"""
import math

import numpy as np

from .pipeline import SynthConfig, thresholds_L


def config1(seed=1, scale=1.0):
    G = int(1_000_000 * scale)
    return dict(k=31, thresholds=thresholds_L(3),
                synth=SynthConfig(seed=seed, genome_len=G, n_reads=int(round(30 * G / 100)),
                                  read_len=100, sub_rate=0.005))


def config2(seed=2, scale=1.0):
    G = int(100_000_000 * scale)
    return dict(k=31, thresholds=thresholds_L(3),
                synth=SynthConfig(seed=seed, genome_len=G, n_reads=int(round(50 * G / 150)),
                                  read_len=150, sub_rate=0.01, n_families=50,
                                  family_min_len=500, family_max_len=5000,
                                  repeat_fraction=0.02, family_divergence=0.01))


def config3(seed=3, scale=1.0):
    G = int(20_000_000 * scale)
    return dict(k=63, thresholds=thresholds_L(3),
                synth=SynthConfig(seed=seed, genome_len=G, n_reads=int(round(40 * G / 250)),
                                  read_len=250, sub_rate=0.005, indel_rate=0.0005,
                                  n_families=50, family_min_len=500, family_max_len=5000,
                                  repeat_fraction=0.02, family_divergence=0.01))


def config4_jobs(seed=4, n_jobs=256, scale=1.0, k=31):
    """The job list of the batched workload (seeded, fixed)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        G = int(round(math.exp(rng.uniform(math.log(50_000), math.log(20_000_000))) * scale))
        G = max(G, 1000)
        L = int(rng.choice([100, 150, 250]))
        cov = float(rng.uniform(20, 60))
        jobs.append(SynthConfig(seed=seed * 100003 + i, genome_len=G, n_reads=max(1, int(round(cov * G / L))),
                                read_len=L, sub_rate=0.01))
    return jobs


def config5_jobs(seed=5, n_jobs=256, rate=40.0, scale=1.0):
    """The served workload: jobs from the config4 distribution, k = 31 or 63
    (80/20), arriving as a Poisson process at `rate` jobs per second.
    Returns (job list, k per job, arrival times in seconds)."""
    cfgs = config4_jobs(seed=seed, n_jobs=n_jobs, scale=scale)
    rng = np.random.default_rng(seed + 1)
    ks = [31 if x < 0.8 else 63 for x in rng.random(n_jobs)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    return cfgs, ks, arrivals


def config4(seed=4, scale=1.0, n_jobs=256):
    return dict(k=31, thresholds=thresholds_L(3), jobs=config4_jobs(seed, n_jobs, scale))


CONFIGS = {"config1": config1, "config2": config2, "config3": config3, "config4": config4}
