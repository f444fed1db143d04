"""Flame statistics from the per-step diagnostics series (FireSim.diag_series()).

The reference plans these as its physics calibration tests (research/DESIGN.md:431-433,
research/physics.md:811-819) without implementing them:
  * flame height: the height where the flame's intermittency on the axis falls to 0.5
    (intermittency I(z) = fraction of time the axis at height z holds flame, i.e. a
    temperature above the flame threshold);
  * puffing frequency: the dominant peak of the spectrum of the heat release Q(t)
    (physics.md allows Q in place of the total visible emission), to be compared with
    f = 1.5 / sqrt(D) Hz (README.md:15, physics.md:355-358).
"""
import numpy as np


def intermittency(axisT, t_flame=1000.0):
    """I(z) per y-plane: fraction of records whose axis temperature exceeds t_flame."""
    a = np.asarray(axisT, np.float64)
    return (a > t_flame).mean(axis=0)


def flame_height(axisT, dx, t_flame=1000.0, level=0.5):
    """Height (m) at which the axis intermittency last falls below `level` going up: the top
    plane with I >= level and the next one above, linearly interpolated between plane centres
    ((j + 0.5) dx).  0.0 if no plane reaches `level`; the domain top if the top plane does."""
    I = intermittency(axisT, t_flame)
    z = (np.arange(I.size) + 0.5) * dx
    above = np.nonzero(I >= level)[0]
    if above.size == 0:
        return 0.0
    j = above[-1]
    if j == I.size - 1:
        return float(I.size * dx)
    return float(z[j] + (I[j] - level) / (I[j] - I[j + 1]) * dx)


def tip_height(tip):
    """Median of the instantaneous flame tip (the 50 % point of the tip's own intermittency)."""
    return float(np.median(np.asarray(tip, np.float64)))


def spectrum(series, dt, segment=None):
    """One-sided power spectrum by Welch's method: linear-detrended, Hann-windowed segments of
    `segment` seconds (default: the whole record if it is under 8 s, else 4 s) with 50 %
    overlap, averaged.  Returns (frequencies, power)."""
    x = np.asarray(series, np.float64)
    n = x.size
    T = n * dt
    seg = n if (segment is None and T < 8.0) else int(round((segment or 4.0) / dt))
    seg = max(8, min(seg, n))
    step = max(1, seg // 2)
    w = np.hanning(seg)
    t = np.arange(seg)
    acc = None
    count = 0
    for s0 in range(0, n - seg + 1, step):
        y = x[s0:s0 + seg]
        y = y - np.polyval(np.polyfit(t, y, 1), t)
        p = np.abs(np.fft.rfft(y * w)) ** 2
        acc = p if acc is None else acc + p
        count += 1
    return np.fft.rfftfreq(seg, dt), acc / count


def puffing_frequency(series, dt, fmin=0.75, fmax=20.0):
    """Dominant frequency (Hz) of the series between fmin (at least two periods over the
    record) and fmax, refined by a parabola through the peak bin and its neighbours.  The
    default band starts at 0.75 Hz: buoyant puffing of fires from 0.1 to 4 m in diameter
    lies above it (1.5 / sqrt(D)), slower drift of the burning rate below."""
    f, p = spectrum(series, dt)
    T = (f.size - 1) * 2 * dt              # the analysed segment length
    lo = max(fmin if fmin is not None else 0.0, 2.0 / T)
    sel = np.nonzero((f >= lo) & (f <= fmax))[0]
    if sel.size < 3:
        raise ValueError("series too short for the requested band")
    k = sel[np.argmax(p[sel])]
    if 0 < k < len(p) - 1:
        a, b, c = np.log(p[k - 1] + 1e-300), np.log(p[k] + 1e-300), np.log(p[k + 1] + 1e-300)
        den = a - 2 * b + c
        off = 0.5 * (a - c) / den if den != 0 else 0.0
        return float(f[k] + np.clip(off, -0.5, 0.5) * (f[1] - f[0]))
    return float(f[k])


def puffing_correlation(D):
    """f = 1.5 / sqrt(D) Hz, D the fire diameter in metres (README.md:15)."""
    return 1.5 / np.sqrt(D)


def summary(series, dx, dt, t_flame=1000.0, skip=0):
    """Flame statistics of a diag_series() dict, dropping the first `skip` records (start-up)."""
    s = {k: np.asarray(v)[skip:] for k, v in series.items()}
    out = {"records": int(s["time"].size), "Q_kW_mean": float(s["Q"].mean() / 1e3),
           "Q_kW_std": float(s["Q"].std() / 1e3), "flame_height_m": flame_height(s["axisT"], dx, t_flame),
           "tip_median_m": tip_height(s["tip"]), "Tmax_mean_K": float(s["Tmax"].mean())}
    for key, name in (("Q", "puffing_Hz_Q"), ("tip", "puffing_Hz_tip")):
        try:
            out[name] = puffing_frequency(s[key], dt)
        except ValueError:
            out[name] = None
    return out
