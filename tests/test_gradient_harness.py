"""The gradient harness's random initial field (hoc.gradient.gaussian_field): a zero-mean Gaussian
field sampled by random Fourier features, whose covariance must be amp^2 exp(-r^2 / (2 corr^2)).
Checked by Monte Carlo over seeds: the sample variance at a point and the sample correlation at
distances of one and two correlation lengths."""
import numpy as np

from hoc.gradient import gaussian_field


def test_gaussian_field_covariance():
    corr, amp = 0.1, 0.1
    pts = np.array([[0.3, 0.4], [0.4, 0.4], [0.3, 0.6]])  # distances 0.1 and 0.2 from the first point
    samples = np.array([gaussian_field(pts, corr=corr, amp=amp, seed=s, n_feat=400) for s in range(400)])
    var = samples[:, 0].var()
    c1 = np.mean(samples[:, 0] * samples[:, 1]) / var
    c2 = np.mean(samples[:, 0] * samples[:, 2]) / var
    print("variance", var, "corr(0.1)", c1, "corr(0.2)", c2)
    assert abs(samples.mean()) < 0.1 * amp
    assert abs(var / amp**2 - 1.0) < 0.2
    assert abs(c1 - np.exp(-0.5)) < 0.12 and abs(c2 - np.exp(-2.0)) < 0.12
