"""Seeded synthetic scan inputs (torch; CPU or GPU).

Genotype dosages: per SNP MAF ~ U(0.05, 0.5); dosage = Binomial(2, MAF)
+ N(0, 0.05) noise, clipped to [0, 2].  Covariates N(0, 1) plus an
intercept column.  Phenotypes N(0, 1) plus planted effects.  Missing
phenotypes are NaN.

Layouts match the device interface of the C ABI: G is (source columns, n)
and Y is (traits, n), one contiguous row per column of the n-row matrix.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

SEED_BASE = 0x0A0A00


@dataclass
class ScanData:
    XL: np.ndarray                 # (n, l) float64, intercept first
    G: torch.Tensor                # (n_src, n) dosages
    Y: torch.Tensor                # (t, n) phenotypes, NaN = missing
    r: int
    E: Optional[np.ndarray] = None       # (n, n_env) interaction covariates
    col_src: Optional[np.ndarray] = None  # (m*r,) source column per output column
    col_env: Optional[np.ndarray] = None  # (m*r,) env column or -1
    planted: Optional[np.ndarray] = None  # (k, 2) (snp, trait)
    name: str = ""

    @property
    def n(self):
        return self.XL.shape[0]

    @property
    def l(self):
        return self.XL.shape[1]

    @property
    def t(self):
        return self.Y.shape[0]

    @property
    def m(self):
        return (self.col_src.size if self.col_src is not None else self.G.shape[0]) // self.r

    def xr_matrix(self, snps=None) -> np.ndarray:
        """(n, r*m') float64 genotype matrix with interaction columns
        materialised, NaN kept (what the reference's I/O thread hands to the
        scan, src/AIOwrapper.cpp:767-840)."""
        snps = np.arange(self.m) if snps is None else np.asarray(snps)
        js = (snps[:, None] * self.r + np.arange(self.r)[None, :]).reshape(-1)
        srcs = np.asarray(self.col_src)[js] if self.col_src is not None else js
        # only the source columns used, gathered where G lives
        uniq, inv = np.unique(srcs, return_inverse=True)
        G = self.G[torch.as_tensor(uniq, device=self.G.device)].detach().cpu().double().numpy()
        cols = []
        for q, j in enumerate(js):
            v = G[inv[q]].copy()
            if self.col_env is not None and self.col_env[j] >= 0:
                v = v * self.E[:, self.col_env[j]]
            cols.append(v)
        return np.stack(cols, axis=1) if cols else np.zeros((self.n, 0))

    def y_matrix(self, traits=None) -> np.ndarray:
        Y = self.Y.detach().cpu().double().numpy()
        if traits is not None:
            Y = Y[np.asarray(traits)]
        return np.ascontiguousarray(Y.T)


def dosages(m, n, gen, device, dtype=torch.float32, chunk=8192):
    """(m, n) dosage matrix, generated in SNP chunks to bound temporaries."""
    out = torch.empty((m, n), dtype=dtype, device=device)
    maf = torch.rand(m, generator=gen, device=device, dtype=torch.float64) * 0.45 + 0.05
    for s0 in range(0, m, chunk):
        s1 = min(m, s0 + chunk)
        f = maf[s0:s1, None].float()
        u1 = torch.rand((s1 - s0, n), generator=gen, device=device)
        u2 = torch.rand((s1 - s0, n), generator=gen, device=device)
        noise = torch.randn((s1 - s0, n), generator=gen, device=device) * 0.05
        d = (u1 < f).float() + (u2 < f).float() + noise
        out[s0:s1] = d.clamp_(0.0, 2.0).to(dtype)
        del u1, u2, noise, d
    return out


def make(n, l, m, t, *, r=1, interaction=False, miss="none", miss_rate=0.05,
         pattern_groups=8, pattern_rate=0.02, planted=20, beta=0.2, seed=0,
         device="cpu", g_dtype=torch.float32, y_dtype=torch.float64, geno_nan=0.0,
         name="") -> ScanData:
    """General generator.

    miss: "none" | "mcar" (independent per trait, rate miss_rate) |
    "groups" (pattern_groups contiguous trait blocks, each blanking a random
    pattern_rate row subset) | "mixed" (some shared, some distinct, some
    complete traits).  interaction=True gives r=2 columns per SNP ordered
    [E*G, G] (single interaction set with the SNP column kept,
    src/AIOwrapper.cpp:819-838)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(SEED_BASE + seed)
    rng = np.random.default_rng(SEED_BASE + seed)
    XL = np.column_stack([np.ones(n), rng.standard_normal((n, l - 1))])
    E = None
    col_src = col_env = None
    if interaction:
        r = 2
        E = rng.standard_normal((n, 1))
        G = dosages(m, n, gen, device, g_dtype)
        col_src = np.repeat(np.arange(m, dtype=np.int32), 2)
        col_env = np.tile(np.array([0, -1], dtype=np.int32), m)
    else:
        G = dosages(m * r, n, gen, device, g_dtype)
    if geno_nan > 0:
        mask = torch.rand(G.shape, generator=gen, device=device) < geno_nan
        G[mask] = float("nan")
    Y = torch.randn((t, n), generator=gen, device=device, dtype=torch.float64)
    pl = None
    if planted and m > 0 and t > 0:
        k = min(planted, m * t)
        idx = rng.choice(m * t, size=k, replace=False)
        pl = np.stack([idx // t, idx % t], axis=1)
        for s, y in pl:
            src = s * r if col_src is None else col_src[s * r + (1 if interaction else 0)]
            g = G[int(src)].double()
            g = torch.nan_to_num(g, nan=0.0)
            Y[int(y)] += beta * g
    if miss == "mcar":
        mask = torch.rand((t, n), generator=gen, device=device) < miss_rate
        Y[mask] = float("nan")
    elif miss == "groups":
        per = max(1, t // pattern_groups)
        for q in range(pattern_groups):
            y0, y1 = q * per, (t if q == pattern_groups - 1 else (q + 1) * per)
            if y0 >= t:
                break
            rows = rng.choice(n, size=max(1, int(round(pattern_rate * n))), replace=False)
            Y[y0:y1, torch.as_tensor(rows, device=device)] = float("nan")
    elif miss == "mixed":
        shared = rng.choice(n, size=max(1, n // 20), replace=False)
        for y in range(t):
            kind = y % 4
            if kind == 1:
                Y[y, torch.as_tensor(shared, device=device)] = float("nan")
            elif kind == 2:
                rows = rng.choice(n, size=max(1, int(miss_rate * n)), replace=False)
                Y[y, torch.as_tensor(rows, device=device)] = float("nan")
            elif kind == 3 and y % 8 == 3:
                Y[y, torch.as_tensor(shared[: len(shared) // 2], device=device)] = float("nan")
    return ScanData(XL=XL, G=G, Y=Y.to(y_dtype), r=r, E=E, col_src=col_src, col_env=col_env,
                    planted=pl, name=name)


# The workloads of the benchmark configurations (sizes as specified).
CONFIGS = {
    0: dict(n=1000, l=3, m=10_000, t=10, miss="none", planted=20, g_dtype=torch.float32),
    1: dict(n=5000, l=6, m=100_000, t=1000, miss="mcar", miss_rate=0.05, planted=50,
            g_dtype=torch.float64, y_dtype=torch.float64),
    2: dict(n=10_000, l=5, m=200_000, t=2000, interaction=True, miss="groups",
            pattern_groups=8, pattern_rate=0.02, planted=50, g_dtype=torch.float32),
}


def config(idx: int, device="cpu", **over) -> ScanData:
    kw = dict(CONFIGS[idx])
    kw.update(over)
    return make(seed=idx, device=device, name="C%d" % idx, **kw)


# Request workloads of the batched-jobs and serving configurations: one
# float16 dosage panel resident on the device and a stream of scan requests
# with their own trait count, covariates, individual subset and missingness.
SERVE_CONFIGS = {
    # 64 jobs submitted together: t log-uniform [1, 4096], 2-16 covariates
    # (+ intercept), individual subset of U[2000, 20000] rows, MCAR U[0, 5%]
    3: dict(n=20_000, m=500_000, jobs=64, t_max=4096, subset=(2000, 20000), rate=None),
    # Poisson arrivals: t log-uniform [1, 8192], 2-16 covariates, all
    # individuals, MCAR U[0, 5%]
    4: dict(n=20_000, m=1_000_000, jobs=32, t_max=8192, subset=None, rate=1.0),
}


@dataclass
class JobSpec:
    t: int
    l: int
    keep: Optional[np.ndarray]
    miss_rate: float
    arrival: float          # seconds after the first request
    seed: int

    def arrays(self, n):
        """(Y (n, t), XL (n, l)) float64 host arrays, NaN = missing."""
        rng = np.random.default_rng(self.seed)
        XL = np.column_stack([np.ones(n), rng.standard_normal((n, self.l - 1))])
        Y = rng.standard_normal((n, self.t))
        if self.miss_rate > 0:
            Y[rng.random((n, self.t)) < self.miss_rate] = np.nan
        return Y, XL


def job_specs(idx: int, jobs=None, seed=0):
    cfg = SERVE_CONFIGS[idx]
    n = cfg["n"]
    rng = np.random.default_rng(SEED_BASE + 100 * idx + seed)
    k = jobs or cfg["jobs"]
    specs, clock = [], 0.0
    for j in range(k):
        t = int(round(np.exp(rng.uniform(0.0, np.log(cfg["t_max"])))))
        l = int(rng.integers(2, 17)) + 1
        keep = None
        if cfg["subset"] is not None:
            size = int(rng.integers(cfg["subset"][0], cfg["subset"][1] + 1))
            keep = np.zeros(n, dtype=np.uint8)
            keep[rng.choice(n, size=size, replace=False)] = 1
        if cfg["rate"]:
            clock += float(rng.exponential(1.0 / cfg["rate"])) if j else 0.0
        specs.append(JobSpec(t=t, l=l, keep=keep, miss_rate=float(rng.uniform(0.0, 0.05)), arrival=clock,
                             seed=SEED_BASE + 1000 * idx + j))
    return specs


def panel(idx: int, device="cuda", m=None, seed=0):
    """(m, n) float16 dosage panel of a request configuration."""
    cfg = SERVE_CONFIGS[idx]
    gen = torch.Generator(device=device)
    gen.manual_seed(SEED_BASE + 100 * idx + seed)
    return dosages(m or cfg["m"], cfg["n"], gen, device, dtype=torch.float16)
