"""Fused detector chain (spsim_hip_detector_chain) against the oracle: exact
deterministic stages, identical random draws for the same Philox counters,
and distribution tests of the noise."""
import math

import numpy as np
import pytest

import oracle
from spsim_amd import device as dv

pytestmark = pytest.mark.gpu


def params(nx=64, ny=48, bx=4, by=4, pol=0, **kw):
    pw = 2e-5
    p = dict(nx=nx, ny=ny, bin_x=bx, bin_y=by, width=float(np.float32(nx * pw)), height=float(np.float32(ny * pw)),
             distance=0.05, pixel_width=pw, pixel_height=pw, center_x=3e-5, center_y=-1e-5, spherical=0,
             polarization=pol, beam_intensity=1e25, quantum_efficiency=0.15, wavelength=1.3e-9,
             electron_hole_production_energy=5.8e-19, dark_current=0.1, exposure_time=1e-13, readout_noise=10.0,
             linear_full_well=2e5, maximum_value=65535.0)
    p.update(kw)
    return p


def run_chain(gpu, p, I, seed=3, base=0, npatterns=1):
    import torch

    det = dv.detector_params(**p)
    It = torch.from_numpy(np.ascontiguousarray(I, np.float32)).to(gpu)
    out = dv.detector_chain(det, It, npatterns, seed, base,
                            outputs=("thomson", "solid_angle", "photons", "count", "electrons", "real_output",
                                     "noiseless_output"))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().reshape(npatterns, -1) for k, v in out.items()}


@pytest.mark.parametrize("pol", [0, 1, 2, 3])
def test_geometry_stages(gpu, pol):
    p = params(pol=pol)
    I = np.ones(p["nx"] * p["ny"], np.float32)
    o = run_chain(gpu, p, I)
    th = oracle.thomson(p["nx"], p["ny"], p["width"], p["height"], p["distance"], p["center_x"], p["center_y"], pol)
    om = oracle.solid_angle(p["nx"], p["ny"], p["width"], p["height"], p["distance"], p["pixel_width"],
                            p["pixel_height"], p["center_x"], p["center_y"], 0)
    # FP64 libm (asin, cos, atan) on host and device may differ in the last
    # FP64 bit, which can move the FP32 result by one ulp.
    np.testing.assert_array_max_ulp(o["thomson"][0], th, maxulp=1)
    np.testing.assert_array_max_ulp(o["solid_angle"][0], om, maxulp=1)
    assert np.mean(o["solid_angle"][0] == om) > 0.95


@pytest.mark.parametrize("bins", [(4, 4), (1, 1), (8, 8)])
def test_deterministic_stages_exact(gpu, bins):
    """Bins of up to 32 raw pixels run the kernel that keeps the geometry in LDS,
    larger ones (8 x 8) the one-pattern-per-lane kernel."""
    bx, by = bins
    p = params(bx=bx, by=by)
    rng = np.random.default_rng(0)
    I = rng.gamma(0.5, 2e6, p["nx"] * p["ny"]).astype(np.float32)
    o = run_chain(gpu, p, I)
    th = oracle.thomson(p["nx"], p["ny"], p["width"], p["height"], p["distance"], p["center_x"], p["center_y"], 0)
    assert np.array_equal(o["thomson"][0], th)
    ph = oracle.photons(o["thomson"][0], o["solid_angle"][0], I, p["beam_intensity"])
    assert np.array_equal(o["photons"][0], ph)
    epp = oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"])
    np.testing.assert_array_equal(o["electrons"][0], o["count"][0] * np.float32(epp))
    epp_nl = oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"], noiseless=True)
    nl = oracle.noiseless_output(ph, p["nx"], p["ny"], bx, by, p["quantum_efficiency"], epp_nl,
                                 p["linear_full_well"], p["maximum_value"])
    assert np.array_equal(o["noiseless_output"][0], nl)


@pytest.mark.parametrize("bins", [(2, 2), (8, 8)])
def test_many_patterns_match_one_at_a_time(gpu, bins):
    """One call over 37 patterns gives bitwise the outputs of 37 one-pattern calls
    (pattern_base + p). With 2 x 2 bins on 256 x 256 pixels each lane takes its bin
    through runs of 2 patterns, the last run holding one."""
    bx, by = bins
    n, npat = 256, 37
    p = params(nx=n, ny=n, bx=bx, by=by, beam_intensity=1e30)
    rng = np.random.default_rng(17)
    I = (10.0 ** rng.uniform(3.0, 7.5, (npat, n * n))).astype(np.float32)
    seed, base = 11, 5
    o = run_chain(gpu, p, I.ravel(), seed=seed, base=base, npatterns=npat)
    for q in (0, 1, 2, 18, 35, 36):
        o1 = run_chain(gpu, p, I[q], seed=seed, base=base + q)
        for k in o1:
            assert np.array_equal(o[k][q], o1[k][0]), (k, q)


def test_random_draws_match_oracle_counters(gpu):
    p = params(nx=96, ny=80, beam_intensity=1e32)
    rng = np.random.default_rng(1)
    # intensities spanning the three Poisson branches (mean < 10, < 1000, >= 1000):
    # mean = QE r0^2 Omega beam I ~ 1.9e-5 I here
    I = (10.0 ** rng.uniform(4.2, 8.6, p["nx"] * p["ny"])).astype(np.float32)
    seed, base = 99, 7
    o = run_chain(gpu, p, I, seed=seed, base=base)
    lam = (o["photons"][0] * np.float32(p["quantum_efficiency"])).astype(np.float64)
    assert lam.min() < 5 and lam.max() > 2000
    ref = np.array([oracle.poisson(l, seed, i, base) for i, l in enumerate(lam)])
    same = np.mean(o["count"][0] == ref)
    assert same > 0.999, same
    # readout noise per binned pixel from the same counters
    nb = (p["nx"] // 4) * (p["ny"] // 4)
    noise = np.array([p["readout_noise"] * oracle.std_normal(seed, b, base, 2) for b in range(nb)], np.float32)
    ro = oracle.real_output(o["electrons"][0], p["nx"], p["ny"], 4, 4, p["dark_current"], p["exposure_time"], noise,
                            p["linear_full_well"], p["maximum_value"])
    np.testing.assert_allclose(o["real_output"][0], ro, rtol=1e-6, atol=1e-3)


@pytest.mark.parametrize("mean", [0.4, 6.0, 40.0, 700.0, 4000.0])
def test_poisson_moments_on_gpu(gpu, mean):
    p = params(nx=256, ny=256, bx=1, by=1)
    o0 = run_chain(gpu, p, np.ones(256 * 256, np.float32))
    ph1 = o0["photons"][0].astype(np.float64)
    # choose I so that photons * QE is close to `mean` in every pixel
    I = (mean / (ph1 * p["quantum_efficiency"])).astype(np.float32)
    o = run_chain(gpu, p, I, seed=5)
    lam = (o["photons"][0] * np.float32(p["quantum_efficiency"])).astype(np.float64)
    x = o["count"][0].astype(np.float64)
    n = x.size
    if mean < 1000:
        m, v = lam, lam
    else:
        m, v = lam - 0.5, lam + 1.0 / 12
    z = (x - m).sum() / math.sqrt(v.sum())
    assert abs(z) < 5, z
    var_ratio = ((x - m) ** 2).sum() / v.sum()
    assert abs(var_ratio - 1) < 5 * math.sqrt(2.0 / n) * 1.5, var_ratio


def test_readout_noise_and_clamp(gpu):
    p = params(nx=512, ny=512, bx=1, by=1, beam_intensity=0.0, dark_current=0.0, linear_full_well=0.0)
    o = run_chain(gpu, p, np.zeros(512 * 512, np.float32), seed=8)
    r = o["real_output"][0].astype(np.float64)
    s = p["readout_noise"]
    # max(N(0, s), 0): mean s/sqrt(2 pi), P(0) = 1/2, second moment s^2/2
    n = r.size
    assert abs(np.mean(r == 0) - 0.5) < 5 * math.sqrt(0.25 / n)
    assert abs(r.mean() - s / math.sqrt(2 * math.pi)) < 5 * s * math.sqrt(0.5 / n)
    assert abs((r ** 2).mean() - s * s / 2) < 5 * s * s * math.sqrt(2.0 / n)


def test_patterns_and_seeds_are_independent(gpu):
    p = params(nx=128, ny=128, bx=1, by=1, beam_intensity=1e32)
    I = np.full(128 * 128, 3e7, np.float32)  # mean ~ 570 photons
    o = run_chain(gpu, p, np.concatenate([I, I]), seed=1, base=0, npatterns=2)
    a, b = o["count"][0], o["count"][1]
    assert not np.array_equal(a, b)
    o2 = run_chain(gpu, p, I, seed=1, base=1)
    assert np.array_equal(o2["count"][0], b)  # pattern index = pattern_base + p
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.05


def test_ragged_binning_raw_outputs(gpu):
    p = params(nx=30, ny=21, bx=4, by=4)
    I = np.full(30 * 21, 2e6, np.float32)
    o = run_chain(gpu, p, I)
    th = oracle.thomson(30, 21, p["width"], p["height"], p["distance"], p["center_x"], p["center_y"], 0)
    assert np.array_equal(o["thomson"][0], th)  # every raw pixel, including those outside full bins
    assert o["real_output"].shape[1] == (30 // 4) * (21 // 4)


@pytest.mark.parametrize("shape,bins", [((16, 12, 6), (4, 4, 2)), ((15, 10, 7), (2, 3, 3))])
def test_three_d_volume_chain(gpu, shape, bins):
    """3-D detector (src/amplification.c:88-110, 133-150): Thomson and solid angle
    per (x, y) repeated over z, Poisson counters over the flat voxel index, bins
    summed in (z, y, x) order; ragged volumes still give every raw voxel."""
    nx, ny, nz = shape
    bx, by, bz = bins
    p = params(nx=nx, ny=ny, bx=bx, by=by, nz=nz, bin_z=bz, spherical=1, beam_intensity=1e32)
    rng = np.random.default_rng(12)
    I = (10.0 ** rng.uniform(4.0, 8.0, nx * ny * nz)).astype(np.float32)
    seed, base = 31, 2
    o = run_chain(gpu, p, I, seed=seed, base=base)
    th2 = oracle.thomson(nx, ny, p["width"], p["height"], p["distance"], p["center_x"], p["center_y"], 0)
    om2 = oracle.solid_angle(nx, ny, p["width"], p["height"], p["distance"], p["pixel_width"], p["pixel_height"],
                             p["center_x"], p["center_y"], 1)
    assert np.array_equal(o["thomson"][0], np.tile(th2, nz))
    assert np.array_equal(o["solid_angle"][0], np.tile(om2, nz))
    ph = oracle.photons(o["thomson"][0], o["solid_angle"][0], I, p["beam_intensity"])
    assert np.array_equal(o["photons"][0], ph)
    lam = (ph * np.float32(p["quantum_efficiency"])).astype(np.float64)
    ref = np.array([oracle.poisson(l, seed, i, base) for i, l in enumerate(lam)])
    assert np.mean(o["count"][0] == ref) > 0.999
    epp_nl = oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"], noiseless=True)
    nl = oracle.noiseless_output(ph, nx, ny, bx, by, p["quantum_efficiency"], epp_nl, p["linear_full_well"],
                                 p["maximum_value"], nz=nz, bz=bz)
    assert np.array_equal(o["noiseless_output"][0], nl)
    nb = (nx // bx) * (ny // by) * (nz // bz)
    noise = np.array([p["readout_noise"] * oracle.std_normal(seed, b, base, 2) for b in range(nb)], np.float32)
    ro = oracle.real_output(o["electrons"][0], nx, ny, bx, by, p["dark_current"], p["exposure_time"], noise,
                            p["linear_full_well"], p["maximum_value"], nz=nz, bz=bz)
    np.testing.assert_allclose(o["real_output"][0], ro, rtol=1e-6, atol=1e-3)


def run_binned(gpu, p, I, seed=3, base=0, npatterns=1):
    import torch

    det = dv.detector_params(**p)
    It = torch.from_numpy(np.ascontiguousarray(I, np.float32)).to(gpu)
    out = dv.detector_chain_binned(det, It, npatterns, seed, base)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().reshape(npatterns, -1) for k, v in out.items()}


@pytest.mark.parametrize("shape,bins", [((96, 80, 1), (4, 4, 1)),    # 16-byte row loads
                                        ((30, 21, 1), (4, 4, 1)),    # ragged, nx not a multiple of 4
                                        ((96, 80, 1), (8, 16, 1)),   # 128 raw pixels per bin: geometry not in LDS
                                        ((16, 12, 6), (4, 4, 2))])   # 3-D volume
def test_bin_draws_match_oracle(gpu, shape, bins):
    """spsim_hip_detector_chain_binned against the oracle's restatement with the
    same counters, over intensities spanning the three Poisson branches; its
    noiseless output bitwise equals the per-pixel chain's."""
    nx, ny, nz = shape
    bx, by, bz = bins
    p = params(nx=nx, ny=ny, bx=bx, by=by, nz=nz, bin_z=bz, spherical=int(nz > 1), beam_intensity=1e32)
    rng = np.random.default_rng(19)
    npat = 3
    I = (10.0 ** rng.uniform(4.2, 8.6, (npat, nx * ny * nz))).astype(np.float32)
    seed, base = 41, 9
    o = run_binned(gpu, p, I.ravel(), seed=seed, base=base, npatterns=npat)
    full = run_chain(gpu, p, I.ravel(), seed=seed, base=base, npatterns=npat)
    epp = oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"])
    for q in range(npat):
        assert np.array_equal(o["noiseless_output"][q], full["noiseless_output"][q])
        ref = oracle.real_output_bin_draws(full["photons"][q], nx, ny, bx, by, p["quantum_efficiency"], epp,
                                           p["dark_current"], p["exposure_time"], p["readout_noise"],
                                           p["linear_full_well"], p["maximum_value"], seed, base + q, nz=nz, bz=bz)
        close = np.isclose(o["real_output"][q], ref, rtol=1e-6, atol=1e-3)
        assert np.mean(close) > 0.999 or close.size < 1000 and np.all(close), (q, np.mean(close))


def test_bin_draws_patterns_match_one_at_a_time(gpu):
    n, npat = 128, 11
    p = params(nx=n, ny=n, beam_intensity=1e30)
    rng = np.random.default_rng(23)
    I = (10.0 ** rng.uniform(3.0, 7.5, (npat, n * n))).astype(np.float32)
    o = run_binned(gpu, p, I.ravel(), seed=5, base=2, npatterns=npat)
    for q in (0, 5, 10):
        o1 = run_binned(gpu, p, I[q], seed=5, base=2 + q)
        for k in o1:
            assert np.array_equal(o[k][q], o1[k][0]), (k, q)


@pytest.mark.parametrize("bin_mean", [0.2, 3.2, 40.0, 9000.0])
def test_bin_draws_moments_on_gpu(gpu, bin_mean):
    """Joint counts recovered from the real output (no dark current, no readout
    noise, unit ADC) have the moments of Poisson(bin mean)."""
    p = params(nx=256, ny=256, dark_current=0.0, readout_noise=0.0, linear_full_well=0.0)
    o0 = run_chain(gpu, p, np.ones(256 * 256, np.float32))
    ph1 = o0["photons"][0].astype(np.float64)
    I = (bin_mean / 16.0 / (ph1 * p["quantum_efficiency"])).astype(np.float32)
    full = run_chain(gpu, p, I)
    lam_px = (full["photons"][0] * np.float32(p["quantum_efficiency"])).astype(np.float64)
    lam = lam_px.reshape(64, 4, 64, 4).sum(axis=(1, 3)).ravel()
    o = run_binned(gpu, p, I, seed=8)
    epp = float(oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"]))
    x = np.rint(o["real_output"][0].astype(np.float64) / epp)
    n = x.size
    z = (x - lam).sum() / math.sqrt(lam.sum())
    assert abs(z) < 5, z
    vr = ((x - lam) ** 2).sum() / lam.sum()
    assert abs(vr - 1) < 5 * math.sqrt(2.0 / n) * 1.5, vr
