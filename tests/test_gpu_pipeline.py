"""Full-pipeline configuration at its full size: 256 orientations of one
50,000-atom C/N/O/S molecule (radius 10 nm) on a 1024 x 1024 detector (20 um
pixels at 0.05 m, lambda 1.3 nm), batched structure factor (persistent
kernel, hipGraph), then the fused detector chain with 4 x 4 binning (QE 0.15,
readout noise 10 e-, dark current 0.1, full well 2e5, maximum value 65535,
beam 1e25 photons/m^2; e-h energy 5.8e-19 J and exposure 1e-13 s assumed).

Checked: intensities of 3 patterns against the FP64 oracle; the deterministic
chain stages of those patterns bit for bit against the oracle; their random
draws against the oracle's sampler with the same counters; and, over all
256 x 65,536 binned pixels, the Poisson and readout-noise moments recovered
from the real output (src/noise.c:32-51, src/amplification.c:77-114); the
batched chain draws one count per bin (spsim_hip_detector_chain_binned)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from spsim_amd import _lib
from spsim_amd import device as dv
from spsim_amd import workloads as wl
from tests.conftest import oracle_threads, rel_l2
from tests.test_gpu_batch import make_plan

pytestmark = pytest.mark.gpu

SEED = 5
BIN = 4


def chain_params(det):
    return dict(nx=det.n, ny=det.n, bin_x=BIN, bin_y=BIN, width=det.width, height=det.width, distance=det.distance,
                pixel_width=det.pixel, pixel_height=det.pixel, center_x=0.0, center_y=0.0, spherical=0,
                polarization=0, beam_intensity=1e25, quantum_efficiency=0.15, wavelength=det.wavelength,
                electron_hole_production_energy=5.8e-19, dark_current=0.1, exposure_time=1e-13, readout_noise=10.0,
                linear_full_well=2e5, maximum_value=65535.0)


def bin_sum(a, n):
    return a.reshape(n // BIN, BIN, n // BIN, BIN).sum(axis=(1, 3)).ravel()


@pytest.fixture(scope="module")
def pipeline(gpu):
    import torch

    det, Z, pos, rots = wl.config5()
    n, npat = det.n, len(rots)
    q = oracle.hkl_detector(n, n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength)
    L = _lib.load()
    plan, *_ = make_plan([(Z, pos)], rots, np.zeros(npat, np.int32), q, n, n)
    try:
        stream = torch.cuda.Stream()
        sp = C.c_void_p(stream.cuda_stream)
        I = torch.empty((npat, n * n), dtype=torch.float32, device=gpu)
        with torch.cuda.stream(stream):
            _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I.data_ptr()), 1, sp), "batch_run")
            p = chain_params(det)
            out = dv.detector_chain_binned(dv.detector_params(**p), I, npat, SEED, 0, stream=stream)
        stream.synchronize()
        res = {"det": det, "Z": Z, "pos": pos, "rots": rots, "q": q, "params": p, "I_dev": I,
               "I": I.cpu().numpy(), "real": out["real_output"].cpu().numpy(),
               "noiseless": out["noiseless_output"].cpu().numpy()}
    finally:
        L.spsim_hip_batch_destroy(plan)
    yield res
    del res


def test_intensities_vs_fp64(pipeline):
    q, Z, pos, rots, I = pipeline["q"], pipeline["Z"], pipeline["pos"], pipeline["rots"], pipeline["I"]
    rng = np.random.default_rng(55)
    errs = []
    for p in (0, 128, 255):
        sub = np.sort(rng.choice(q.shape[0], 2048, replace=False))
        _, ref = oracle.sf_f64(q[sub], Z, pos, R=rots[p], threads=oracle_threads())
        errs.append(rel_l2(I[p, sub], ref))
    print(f"config 5: rel err vs FP64 (3 patterns x 2048 px) max {max(errs):.2e}")
    assert max(errs) < 1e-5


def test_chain_stages_of_three_patterns(gpu, pipeline):
    """The per-pixel chain of one pattern on its own (pattern_base = p, all raw
    outputs): its noiseless output equals the batch call's bit for bit; its
    deterministic stages equal the oracle's; its Poisson draws equal the
    oracle sampler's with the same counters (>= 99.9 % of a pixel sample) and
    its readout output the oracle's with the same Gaussian draws. The batch
    call's real output (one count per bin) equals the oracle's restatement of
    the per-bin draws on >= 99.9 % of the bins."""
    det, p, I_dev = pipeline["det"], pipeline["params"], pipeline["I_dev"]
    n = det.n
    th = oracle.thomson(n, n, p["width"], p["height"], p["distance"], 0, 0, 0)
    om = oracle.solid_angle(n, n, p["width"], p["height"], p["distance"], p["pixel_width"], p["pixel_height"])
    epp = oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"])
    epp_nl = oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"], noiseless=True)
    rng = np.random.default_rng(56)
    for pat in (0, 128, 255):
        o = dv.detector_chain(dv.detector_params(**p), I_dev[pat].contiguous(), 1, SEED, pat,
                              outputs=("thomson", "solid_angle", "photons", "count", "electrons", "real_output",
                                       "noiseless_output"))
        o = {k: v.cpu().numpy()[0] for k, v in o.items()}
        assert np.array_equal(o["noiseless_output"], pipeline["noiseless"][pat])
        rb = oracle.real_output_bin_draws(o["photons"], n, n, BIN, BIN, p["quantum_efficiency"], epp,
                                          p["dark_current"], p["exposure_time"], p["readout_noise"],
                                          p["linear_full_well"], p["maximum_value"], SEED, pat)
        assert np.mean(np.isclose(pipeline["real"][pat], rb, rtol=1e-6, atol=1e-3)) > 0.999
        np.testing.assert_array_max_ulp(o["thomson"], th, maxulp=1)
        np.testing.assert_array_max_ulp(o["solid_angle"], om, maxulp=1)
        ph = oracle.photons(o["thomson"], o["solid_angle"], pipeline["I"][pat], p["beam_intensity"])
        assert np.array_equal(o["photons"], ph)
        nl = oracle.noiseless_output(ph, n, n, BIN, BIN, p["quantum_efficiency"], epp_nl, p["linear_full_well"],
                                     p["maximum_value"])
        assert np.array_equal(o["noiseless_output"], nl)
        np.testing.assert_array_equal(o["electrons"], o["count"] * np.float32(epp))
        lam = (ph * np.float32(p["quantum_efficiency"])).astype(np.float64)
        sub = rng.choice(n * n, 20000, replace=False)
        sub = np.union1d(sub, np.argsort(lam)[-2000:])  # the brightest pixels, where draws are not all 0
        ref = np.array([oracle.poisson(lam[i], SEED, int(i), pat) for i in sub])
        assert np.mean(o["count"][sub] == ref) > 0.999
        nb = (n // BIN) ** 2
        noise = np.array([p["readout_noise"] * oracle.std_normal(SEED, b, pat, 2) for b in range(nb)], np.float32)
        ro = oracle.real_output(o["electrons"], n, n, BIN, BIN, p["dark_current"], p["exposure_time"], noise,
                                p["linear_full_well"], p["maximum_value"])
        np.testing.assert_allclose(o["real_output"], ro, rtol=1e-6, atol=1e-3)


def test_noise_moments_over_all_patterns(pipeline):
    """Over all 256 x 65,536 binned pixels: the photon count N_b of each bin is
    recovered from the real output (263 electrons per photon against 10 e-
    readout noise); N_b ~ Poisson(sum of the bin's QE x photons), and the
    residual is N(0, 10 e-) where N_b >= 1 and max(N(0, 10), 0) where N_b = 0."""
    det, p = pipeline["det"], pipeline["params"]
    n = det.n
    th = oracle.thomson(n, n, p["width"], p["height"], p["distance"], 0, 0, 0).astype(np.float64)
    om = oracle.solid_angle(n, n, p["width"], p["height"], p["distance"], p["pixel_width"],
                            p["pixel_height"]).astype(np.float64)
    k = th * om * p["beam_intensity"] * p["quantum_efficiency"]
    epp = float(oracle.electrons_per_photon(p["wavelength"], p["electron_hole_production_energy"]))
    adc = p["maximum_value"] / p["linear_full_well"]
    dark = BIN * BIN * p["dark_current"] * p["exposure_time"]
    I, real = pipeline["I"], pipeline["real"]
    lam = np.stack([bin_sum(k * I[pat].astype(np.float64), n) for pat in range(I.shape[0])])
    e = real.astype(np.float64) / adc - dark
    N = np.rint(e / epp)
    r = e - N * epp
    print(f"config 5 noise: sum lambda {lam.sum():.4g}, bins with >= 1 photon {np.mean(N >= 1):.4f}, "
          f"max lambda {lam.max():.3g}")
    assert lam.sum() > 1e5  # enough photons for the moment tests to mean something
    # Poisson: mean and variance of N - lambda
    z = (N - lam).sum() / math.sqrt(lam.sum())
    assert abs(z) < 5, z
    v = ((N - lam) ** 2).sum()
    sd = math.sqrt((2 * lam ** 2 + lam).sum())
    assert abs(v - lam.sum()) < 5 * sd, (v, lam.sum(), sd)
    # readout noise on bins with photons: N(0, 10)
    s = p["readout_noise"]
    rp = r[N >= 1]
    assert abs(rp.mean()) < 5 * s / math.sqrt(rp.size)
    assert abs((rp ** 2).mean() / (s * s) - 1) < 5 * math.sqrt(2.0 / rp.size)
    # bins without photons: max(N(0, 10), 0): half exactly 0, mean s / sqrt(2 pi)
    r0 = real[N == 0].astype(np.float64) / adc
    assert abs(np.mean(r0 == 0) - 0.5) < 5 * math.sqrt(0.25 / r0.size)
    assert abs(r0.mean() - s / math.sqrt(2 * math.pi)) < 5 * s * math.sqrt(0.5 / r0.size)
