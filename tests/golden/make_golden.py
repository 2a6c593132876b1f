"""Write the golden fixtures of the hot path with an independent numpy FP64
implementation (numpy's pocketfft, plain numpy loops).  No code of the
reference, of the oracle or of the HIP library is used, so the fixtures pin
the oracle's restatement of:

  amplitudes   all_vectors_scatter_device.cpp:418-438,
               self_vectors_scatter_device.cpp:306-319
  correlation  smath.cpp:141-156 (fftw: pad to 2NF, |X|^2, inverse,
               1/(2NF (NF-tau))) and smath.cpp:51-76 (direct, whose
               imaginary part is the negative of the fftw method's)
  dsp/store    all_vectors_scatter_device.cpp:209-236, 355-360,
               self_vectors_scatter_device.cpp:115-142, 233-238

Run:  python tests/golden/make_golden.py   (writes tests/golden/*.npz)
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def amplitudes_all(coords_frames, b, q):
    # coords [NF][NA][3] float32 promoted to double; p = x qx + y qy + z qz
    c = coords_frames.astype(np.float64)
    p = c[:, :, 0] * q[0] + c[:, :, 1] * q[1] + c[:, :, 2] * q[2]
    return (b[None, :] * np.cos(p)).sum(axis=1) + 1j * (b[None, :] * np.sin(p)).sum(axis=1)


def amplitude_self(coords_atoms, j, s, q):
    c = coords_atoms[j].astype(np.float64)
    p = c[:, 0] * q[0] + c[:, 1] * q[1] + c[:, 2] * q[2]
    return s * np.cos(p) + 1j * s * np.sin(p)


def dsp(a, kind, method):
    NF = a.size
    if kind == "autocorrelate":
        if method == "fftw":
            X = np.fft.fft(np.concatenate([a, np.zeros(NF)]))
            R = np.fft.ifft(np.abs(X) ** 2) * (2 * NF)
            return R[:NF] / (2 * NF * (NF - np.arange(NF)))
        # direct: sum_k a_k conj(a_{k+tau}) / (NF - tau)
        return np.array([np.sum(a[:NF - t] * np.conj(a[t:])) / (NF - t) for t in range(NF)])
    if kind == "square":
        return np.abs(a) ** 2 + 0j
    if kind == "plain":
        return a.copy()
    raise ValueError(kind)


def shell(kind, coords, b, qv, dsp_kind, method):
    NM = qv.shape[0]
    if kind == "all":
        NF = coords.shape[0]
        series = [dsp(amplitudes_all(coords, b, q), dsp_kind, method) for q in qv]
    else:
        NA, NF = coords.shape[0], coords.shape[1]
        series = [dsp(amplitude_self(coords, j, b[j], q), dsp_kind, method)
                  for j in range(NA) for q in qv]
    series = np.array(series)
    a = series.mean(axis=1)
    fqt = series.sum(axis=0) / NM
    fq = a.sum() / NM
    fq2 = np.sum(np.abs(a) ** 2) / NM
    return fqt, fq, fq2 + 0j


def make_case(name, kind, NA, NF, NM, qlen, seed, box=6.0, step=0.3):
    rng = np.random.default_rng(seed)
    start = rng.uniform(0.0, box, size=(NA, 3))
    walk = np.cumsum(rng.normal(0.0, step, size=(NF, NA, 3)), axis=0)
    frames = (start[None, :, :] + walk).astype(np.float32)          # [NF][NA][3]
    b = rng.choice([-3.74, 6.65, 9.36, 5.80], size=NA)
    u = rng.normal(size=(NM, 3))
    u /= np.sqrt((u * u).sum(axis=1))[:, None]
    out = {"frames": frames, "b": b, "u": u, "qlen": np.array(qlen, dtype=np.float64)}
    coords = frames if kind == "all" else np.ascontiguousarray(frames.transpose(1, 0, 2))
    for dk, mt in (("autocorrelate", "fftw"), ("autocorrelate", "direct"),
                   ("square", "fftw"), ("plain", "fftw")):
        fqts, fqs, fq2s = [], [], []
        for ql in qlen:
            fqt, fq, fq2 = shell(kind, coords, b, ql * u, dk, mt)
            fqts.append(fqt); fqs.append(fq); fq2s.append(fq2)
        tag = f"{dk}_{mt}"
        out[f"fqt_{tag}"] = np.array(fqts)
        out[f"fq_{tag}"] = np.array(fqs)
        out[f"fq2_{tag}"] = np.array(fq2s)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), kind=kind, **out)


def main():
    make_case("golden_all_small", "all", NA=7, NF=12, NM=3, qlen=[0.5, 1.7], seed=11)
    make_case("golden_all_prime", "all", NA=5, NF=13, NM=4, qlen=[2.3], seed=12)
    make_case("golden_all_oneframe", "all", NA=3, NF=1, NM=2, qlen=[1.0], seed=13)
    make_case("golden_self_small", "self", NA=4, NF=9, NM=2, qlen=[0.8, 3.1], seed=14)
    make_case("golden_self_oneatom", "self", NA=1, NF=17, NM=5, qlen=[1.3], seed=15)


if __name__ == "__main__":
    main()
