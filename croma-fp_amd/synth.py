"""Seeded synthetic tracer workloads for the benchmark configurations.

The reference's own synthetic mode (tracer_fp_debug.c:34-116) is a
deterministic sine timeline; the benchmark configurations instead specify
seeded log-uniform draws, generated here with numpy's PCG64:

  n_gas   ~ logU(1e-5, 1e-2) cm^-3
  T       ~ logU(1e6, 1e8) K           -> kbt = kB T 1e-9 GeV (fp_shared_core.c:920)
  |B|     ~ logU(0.01, 10) uG          -> Gauss
  div v   ~ N(0, 1e-17 s^-1)           -> 1/Gyr (x 3.1536e16)
  l_turb  = 150 kpc (L_turb_target_kpc, params.c:87)
  c_s     = c_sound(T)                  (FP_Coef.c:897-900)
  dv      = curl_v l_turb, curl_v ~ 3e-17 s^-1 x logN(0, sigma), capped at
            mach_limit c_s (dv_limit, FP_Coef.c:614-620)
  beta_pl = n kB T eV / (B^2 / 8 pi)    (fp_shared_core.c:899-903)

Backgrounds drift between snapshots by a log-space random walk (5% per
snapshot in n, T, B; fresh div v draws) so that the coefficient
interpolation between snapshots is exercised.  Redshift runs linearly from
z0 to 0; snapshot durations are flat LCDM age differences (H0 = 70,
Omega_m = 0.3).  The initial spectrum is N(p) = A p^-2.2 exp(-p / p_cut).
Configs with dsa_frac > 0 add shock-injected sources on a random fraction
of (tracer, snapshot) pairs: Mach ~ logU(2, 10), the pre-shock temperature
and density the tracer's background at that snapshot, through the library's
DSA builder (eta(M)-normalised proton and electron spectra per unit volume,
times the tracer's downstream volume, per Gyr; tracer_fp_dsa.c:184-505).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

KB_EV = 8.62e-5
EV_ERG = 1.6e-12
C_LIGHT = 2.999e10
MU_MOL = 0.59
M_PROTON_G = 1.6726e-24
GYR_S = 3.1536e16
MPC_CM = 3.0857e24
FIELDS = ("n_gas", "kbt", "b_field", "divv", "l_turb", "dv", "cs", "beta_pl")


def c_sound(temp):
    return np.sqrt(5.0 / 3.0 * KB_EV * temp / (MU_MOL * 0.938 * 1.0e9)) * C_LIGHT


def lcdm_age_gyr(z, h0=70.0, om=0.3):
    ol = 1.0 - om
    hubble_time_gyr = 977.792 / h0
    return 2.0 / (3.0 * math.sqrt(ol)) * hubble_time_gyr * np.arcsinh(np.sqrt(ol / om) * (1.0 + np.asarray(z)) ** -1.5)


@dataclass
class Workload:
    n: int
    ntracer: int
    nsnap: int
    species: int                  # seed species: 1 electron_only, 2 proton_only, 3 electron_proton
    dpp_model: int                # 1 asa, 2 ttd, 3 off
    z: np.ndarray                 # [nsnap + 1]
    dt_snap: np.ndarray           # [nsnap]
    bg: np.ndarray                # [nsnap + 1, 8, ntracer]
    crp0: np.ndarray              # [ntracer * n]
    cre0: np.ndarray              # [ntracer * n]
    mass: np.ndarray              # [ntracer]
    qpi: dict = field(default_factory=dict)    # snapshot -> [ntracer * n]
    qepri: dict = field(default_factory=dict)  # snapshot -> [ntracer * n]
    pmin: float = -1.0
    pmax: float = 8.0
    pemin: float = 0.0
    pemax: float = 6.0

    @property
    def protons(self):
        return self.species in (2, 3)

    def subset(self, idx):
        """Tracer subset (same snapshots), e.g. for a CPU baseline sample."""
        idx = np.asarray(idx)
        n = self.n
        rows = (idx[:, None] * n + np.arange(n)[None, :]).ravel()
        w = Workload(self.n, len(idx), self.nsnap, self.species, self.dpp_model, self.z, self.dt_snap,
                     np.ascontiguousarray(self.bg[:, :, idx]), self.crp0[rows].copy(), self.cre0[rows].copy(),
                     self.mass[idx].copy(), pmin=self.pmin, pmax=self.pmax, pemin=self.pemin, pemax=self.pemax)
        w.qpi = {s: v[rows].copy() for s, v in self.qpi.items()}
        w.qepri = {s: v[rows].copy() for s, v in self.qepri.items()}
        return w


def grid_axis(n, lo, hi):
    """Bin centres of the reference log grid (FP_Coef.c:291-330), p = 10^(lo + j*step)."""
    step = (hi - lo) / n
    a = lo - step
    p = np.empty(n)
    for j in range(n):
        a = a + step
        p[j] = 10.0 ** a
    return p


def make_workload(n=128, ntracer=1024, nsnap=100, species=1, dpp_model=3, seed=1, z0=1.0, dsa_frac=0.0,
                  curl_sigma=1.0, mach_limit=0.5, e_cut=1.0e4, p_cut=1.0e6, walk=0.05):
    rng = np.random.default_rng(seed)
    nt = ntracer
    z = np.linspace(z0, 0.0, nsnap + 1)
    age = lcdm_age_gyr(z)
    dt = np.diff(age)
    log_n = rng.uniform(-5.0, -2.0, nt)
    log_t = rng.uniform(6.0, 8.0, nt)
    log_b = rng.uniform(-2.0, 1.0, nt)
    log_curl = np.log(3.0e-17) + curl_sigma * rng.standard_normal(nt)
    bg = np.empty((nsnap + 1, 8, nt))
    l_turb_mpc = 0.15
    for s in range(nsnap + 1):
        if s > 0:
            log_n = log_n + walk * rng.standard_normal(nt) / math.log(10.0)
            log_t = np.clip(log_t + walk * rng.standard_normal(nt) / math.log(10.0), 5.5, 8.5)
            log_b = log_b + walk * rng.standard_normal(nt) / math.log(10.0)
            log_curl = log_curl + walk * rng.standard_normal(nt)
        n_gas = 10.0 ** log_n
        temp = 10.0 ** log_t
        b = 10.0 ** log_b * 1.0e-6
        divv = rng.normal(0.0, 1.0e-17, nt) * GYR_S
        cs = c_sound(temp)
        dv = np.exp(log_curl) * l_turb_mpc * MPC_CM
        dv = np.where(dv / cs > mach_limit, mach_limit * cs, dv)
        beta = n_gas * KB_EV * temp * EV_ERG / (b * b / (8.0 * math.pi))
        bg[s] = np.stack([n_gas, KB_EV * temp * 1.0e-9, b, divv, np.full(nt, l_turb_mpc), dv, cs, beta])
    pe = grid_axis(n, 0.0, 6.0)
    pp = grid_axis(n, -1.0, 8.0)
    spec_e = pe ** -2.2 * np.exp(-pe / e_cut)
    cre0 = np.tile(spec_e, nt) * np.repeat(10.0 ** rng.uniform(-0.5, 0.5, nt), n)
    if species == 2:
        cre0 = np.zeros(nt * n)
    if species in (2, 3):
        spec_p = 100.0 * pp ** -2.2 * np.exp(-pp / p_cut)
        crp0 = np.tile(spec_p, nt) * np.repeat(10.0 ** rng.uniform(-0.5, 0.5, nt), n)
    else:
        crp0 = np.zeros(nt * n)
    w = Workload(n, nt, nsnap, species, dpp_model, z, dt, bg, crp0, cre0, np.full(nt, 1.0e7))
    if dsa_frac > 0.0:
        # shock injection through the library's DSA builder (TRACER_SOURCE
        # mode, tracer_fp_dsa.c:456-505): the hit tracers' pre-shock state is
        # the snapshot's background, rho = mu m_p n_gas
        import croma_hip as ch
        dprm = ch.dsa_default_params()
        for s in range(nsnap):
            hit = rng.random(nt) < dsa_frac
            if not hit.any():
                continue
            mach = np.where(hit, 10.0 ** rng.uniform(math.log10(2.0), 1.0, nt), np.nan)
            temp1 = bg[s, 1] / (KB_EV * 1.0e-9)
            rho1 = bg[s, 0] * MU_MOL * M_PROTON_G
            qe = np.zeros(nt * n)
            qp = np.zeros(nt * n)
            ch.dsa_inject(n, dprm, ch.DSA_SOURCE, mach, rho1, w.mass, temp1=temp1, dt_gyr=float(dt[s]), qpi=qp,
                          qepri=qe, pmin=-1.0, pmax=8.0, pemin=0.0, pemax=6.0)
            w.qepri[s] = qe
            if species in (2, 3):
                w.qpi[s] = qp
    return w


CONFIGS = {
    # plain CPU-runnable case: 1,024 electron tracers, 128 bins, 100 snapshots, D_pp off
    "c0": dict(n=128, ntracer=1024, nsnap=100, species=1, dpp_model=3),
    # batched uneven work: 10,000 tracers, electron_proton, 256 bins, ASA, DSA sources on 5%
    "c1": dict(n=256, ntracer=10000, nsnap=100, species=3, dpp_model=1, dsa_frac=0.05),
    # bulk throughput: 1,000,000 electron tracers, 512 bins, TTD, 20 snapshots
    "c2": dict(n=512, ntracer=1000000, nsnap=20, species=1, dpp_model=2),
}
