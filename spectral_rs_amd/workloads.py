"""The benchmark workloads, in reference units (c = 10 cells/omega_pe and
dt = 0.045: 0.45 cells per step and a 10-cell electron skin depth; ions
q = +1, m = 16 listed first, then electrons q = -1, m = 1, as src/lib.rs:124-126;
4 binomial passes; delta = 8 for grids up to 256 cells wide, else 16).

* C2: the plain FP32 shock at 4096 x 1024, 32 ppc per species, gamma0 = 10,
  T = 1e-4 m_e c^2 (electron vth 0.01, ion vth 0.0025): 266,338,304 particles;
* C3: the precision-study shock at 2048 x 512, 64 ppc per species:
  132,120,576 particles;
* C4: 64 independent shock runs drawn with seed 0 from grids
  {128,256,512,1024} x {32,64,128,256}, ppc {4,8,16,32}, gamma0 uniform in
  [2,20], T log-uniform in [1e-5,1e-3];
* C5: the learned-denoiser shock at 4096 x 1024 with 4 ppc per species
  (33,292,288 particles), against plain 64 ppc (532,676,608).
"""
import numpy as np

SHOCK = dict(n_pass=4, dt=0.045, c=10.0)
SPECIES = [(1.0, 16.0, 0.0025), (-1.0, 1.0, 0.01)]  # (q, m, vth) at T = 1e-4: ions then electrons
C2 = dict(SHOCK, size_x=4096, size_y=1024, delta=16, dens=32, gamma_inj=10.0)
C3 = dict(SHOCK, size_x=2048, size_y=512, delta=16, dens=64, gamma_inj=10.0)
C5 = dict(SHOCK, size_x=4096, size_y=1024, delta=16, dens=4, gamma_inj=10.0)


def delta_for(size_x):
    return 8 if size_x <= 256 else 16


def particles(cfg):
    """Both species: 2 ppc (size_x - 2 delta) size_y (src/lib.rs:258-260)."""
    return 2 * cfg["dens"] * (cfg["size_x"] - 2 * cfg["delta"]) * cfg["size_y"]


def ensemble_runs(count=64, seed=0):
    """The C4 draw: `count` run configurations (sp_params keys plus `temp`,
    electron vth = sqrt(T), ion vth = sqrt(T/16))."""
    rng = np.random.default_rng(seed)
    runs = []
    for _ in range(count):
        nx = int(rng.choice([128, 256, 512, 1024]))
        ny = int(rng.choice([32, 64, 128, 256]))
        ppc = int(rng.choice([4, 8, 16, 32]))
        g0 = float(rng.uniform(2.0, 20.0))
        temp = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e-3))))
        runs.append(dict(SHOCK, size_x=nx, size_y=ny, delta=delta_for(nx), dens=ppc, gamma_inj=g0, temp=temp))
    return runs
