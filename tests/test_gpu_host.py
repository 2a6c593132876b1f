"""End-to-end runs of the C++ host layer on the GPU: scatter.xml -> DCD ->
ScatterDeviceFactory -> device -> HDF5 signal file, checked shell by shell
against the FP64 oracle, including the resume of a partly written file."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from hostio import read_raw_signal, sx_run, write_dcd  # noqa: E402
from util import fq2_bound, shell_error  # noqa: E402

pytestmark = pytest.mark.gpu

B = {"H": -3.7390, "C": 6.6460, "N": 9.36, "O": 5.803}


def setup_case(tmp_path, NA=150, NF=120, NM=16, seed=5):
    rng = np.random.default_rng(seed)
    frames = (rng.uniform(0, 25, size=(1, NA, 3)) + np.cumsum(rng.normal(0, 0.15, size=(NF, NA, 3)), axis=0))
    frames = frames.astype(np.float32)
    elements = rng.choice(list(B), size=NA)
    write_dcd(tmp_path / "traj.dcd", frames, unit_cell=True)
    (tmp_path / "atoms.txt").write_text("\n".join(elements) + "\n")
    u = rng.normal(size=(NM, 3))
    (tmp_path / "orient.txt").write_text("\n".join(" ".join(f"{v:.17g}" for v in r) for r in u))
    b = np.array([B[e] for e in elements])
    return frames, b, u / np.linalg.norm(u, axis=1)[:, None]


def config(kind, qfile, orientation="vectors", extra=""):
    orient = {"vectors": "<type>vectors</type><vectors><type>file</type><file>orient.txt</file></vectors>",
              "none": "<type>none</type>",
              "sphere": "<type>multipole</type><multipole><type>sphere</type><moments><type>resolution</type>"
                        "<resolution>5</resolution></moments></multipole>",
              "cylinder": "<type>multipole</type><axis><x>0.2</x><y>0.1</y><z>1</z></axis><multipole>"
                          "<type>cylinder</type><moments><type>resolution</type><resolution>4</resolution>"
                          "</moments></multipole>"}[orientation]
    return f"""<?xml version="1.0"?>
<root>
  <sample>
    <structure><file>atoms.txt</file><format>elements</format></structure>
    <framesets><frameset><file>traj.dcd</file><format>dcd</format></frameset></framesets>
  </sample>
  <scattering>
    <type>{kind}</type>
    <vectors><type>file</type><file>{qfile}</file></vectors>
    <average><orientation>{orient}</orientation></average>
    <dsp><type>autocorrelate</type><method>fftw</method></dsp>
    <signal><file>signal.h5</file></signal>
    {extra}
  </scattering>
</root>"""


def read_signal(tmp_path):
    sx_run("--h5-to-raw", str(tmp_path / "signal.h5"), str(tmp_path / "signal.raw"))
    return read_raw_signal(tmp_path / "signal.raw")


def check_shells(oracle_mod, kind, frames, b, u, q, fqt, fq, fq2):
    coords = frames if kind == "all" else np.ascontiguousarray(frames.transpose(1, 0, 2))
    for i in range(q.shape[0]):
        qv = oracle_mod.init_subvectors(q[i], u)
        r = oracle_mod.compute(kind, coords, b, qv, nthreads=8)
        assert shell_error(fqt[i], r[0])[0] < 1e-5
        assert abs(fq[i] - r[1]) < 1e-5 * abs(r[0][0])
        assert abs(fq2[i] - r[2]) <= fq2_bound(r[2], r[0][0])


@pytest.mark.parametrize("kind", ["all", "self"])
def test_sx_run_dcd_to_hdf5_with_resume(tmp_path, oracle_mod, kind):
    frames, b, u = setup_case(tmp_path)
    qs = np.array([[0.4, 0.0, 0.0], [0.0, 1.1, 0.3], [2.2, -0.5, 0.7], [0.0, 0.0, 3.0], [1.0, 1.0, 1.0]])
    (tmp_path / "q3.txt").write_text("\n".join(" ".join(f"{v:.17g}" for v in r) for r in qs[:3]))
    (tmp_path / "q5.txt").write_text("\n".join(" ".join(f"{v:.17g}" for v in r) for r in qs))
    (tmp_path / "a.xml").write_text(config(kind, "q3.txt"))
    (tmp_path / "b.xml").write_text(config(kind, "q5.txt"))
    sx_run("--config", str(tmp_path / "a.xml"), cwd=tmp_path)
    q, fqt, fq, fq2 = read_signal(tmp_path)
    assert np.array_equal(q, qs[:3])
    res = sx_run("--config", str(tmp_path / "b.xml"), cwd=tmp_path)
    assert "2 of 5 q vectors" in res.stderr
    q, fqt, fq, fq2 = read_signal(tmp_path)
    assert np.array_equal(q, qs)
    check_shells(oracle_mod, kind, frames, b, u, q, fqt, fq, fq2)
    res = sx_run("--config", str(tmp_path / "b.xml"), cwd=tmp_path)
    assert "0 of 5 q vectors" in res.stderr


def test_sx_run_scans_without_orientations(tmp_path, oracle_mod):
    frames, b, u = setup_case(tmp_path, NA=80, NF=90)
    xml = config("all", "unused", orientation="none").replace(
        "<vectors><type>file</type><file>unused</file></vectors>",
        "<vectors><type>scans</type><scans><scan><from>0.2</from><to>2.5</to><points>7</points>"
        "<base><x>0.6</x><y>0.8</y><z>0</z></base></scan></scans></vectors>")
    (tmp_path / "s.xml").write_text(xml)
    sx_run("--config", str(tmp_path / "s.xml"), cwd=tmp_path)
    q, fqt, fq, fq2 = read_signal(tmp_path)
    assert q.shape == (7, 3)
    check_shells(oracle_mod, "all", frames, b, None, q, fqt, fq, fq2)


def test_sx_run_scan_with_orientations_steps_shells(tmp_path, oracle_mod):
    """A scatter.xml q scan (40 points, 0.1-4.0 A^-1, interior points rounded
    by single-precision powf as parameters.cpp:1151 does) with a file of
    orientations: the host hands the shells to the oriented entry point,
    which steps phasors across them with first-order residual corrections
    (reported as "sd:stepped-residual"); every shell against the oracle."""
    frames, b, u = setup_case(tmp_path, NA=200, NF=150, NM=12)
    xml = config("all", "unused").replace(
        "<vectors><type>file</type><file>unused</file></vectors>",
        "<vectors><type>scans</type><scans><scan><from>0.1</from><to>4.0</to><points>40</points>"
        "<base><x>1</x><y>0</y><z>0</z></base></scan></scans></vectors>")
    (tmp_path / "s.xml").write_text(xml)
    res = sx_run("--config", str(tmp_path / "s.xml"), cwd=tmp_path)
    line = next(l for l in res.stderr.splitlines() if "sd:stepped-residual" in l)
    assert int(line.split("count")[1].split()[0]) == 40
    q, fqt, fq, fq2 = read_signal(tmp_path)
    assert np.array_equal(q, oracle_mod.scan(0.1, 4.0, 40))
    check_shells(oracle_mod, "all", frames, b, u, q, fqt, fq, fq2)


@pytest.mark.parametrize("mp", ["sphere", "cylinder"])
def test_sx_run_multipole(tmp_path, oracle_mod, mp):
    """scattering.average.orientation.type=multipole through the factory."""
    frames, b, _ = setup_case(tmp_path, NA=60, NF=90)
    frames = frames - np.float32(12.5)   # about the origin of the expansion
    write_dcd(tmp_path / "traj.dcd", frames, unit_cell=True)
    qs = np.array([[0.3, 0.0, 0.0], [0.0, 0.8, 0.3], [0.5, -0.5, 0.4]])
    (tmp_path / "q.txt").write_text("\n".join(" ".join(f"{v:.17g}" for v in r) for r in qs))
    (tmp_path / "m.xml").write_text(config("all", "q.txt", orientation=mp))
    sx_run("--config", str(tmp_path / "m.xml"), cwd=tmp_path)
    q, fqt, fq, fq2 = read_signal(tmp_path)
    assert np.array_equal(q, qs)
    mom = oracle_mod.sphere_moments(5) if mp == "sphere" else oracle_mod.cylinder_moments(4)
    for i in range(len(qs)):
        r = oracle_mod.compute_multipole(mp, frames, b, qs[i], mom, axis=(0.2, 0.1, 1.0))
        assert shell_error(fqt[i], r[0])[0] < 1e-5
        assert abs(fq[i] - r[1]) < 1e-5 * abs(r[0][0])
        assert abs(fq2[i] - r[2]) <= fq2_bound(r[2], r[0][0])


def test_sx_run_cylinder_vectors_along_a_scan(tmp_path, oracle_mod):
    """Cylinder averaging about an axis over a scan of fixed direction: the
    host hands the shells to the oriented entry point (|q| times one shared
    subvector set); checked against the oracle on the host's subvectors."""
    frames, b, _ = setup_case(tmp_path, NA=70, NF=80)
    scan = ("<vectors><type>scans</type><scans><scan><from>0.3</from><to>2.1</to><points>4</points>"
            "<base><x>0.6</x><y>0.8</y><z>0.3</z></base></scan></scans></vectors>")
    orient = ("<type>vectors</type><axis><x>0.2</x><y>0.1</y><z>1</z></axis><vectors><type>cylinder</type>"
              "<algorithm>raster_linear</algorithm><resolution>30</resolution></vectors>")
    xml = config("all", "unused").replace("<vectors><type>file</type><file>unused</file></vectors>", scan)
    xml = xml.replace("<type>vectors</type><vectors><type>file</type><file>orient.txt</file></vectors>", orient)
    (tmp_path / "c.xml").write_text(xml)
    out = sx_run("--config", str(tmp_path / "c.xml"), "--vectors", cwd=tmp_path).stdout.split("\n")
    k = next(i for i, l in enumerate(out) if l.startswith("subvectors"))
    subs = []
    for l in out[k + 1:k + 1 + int(out[k].split()[1])]:
        v = [float(x) for x in l.split()]
        subs.append(np.array(v[1:]).reshape(int(v[0]), 3))
    sx_run("--config", str(tmp_path / "c.xml"), cwd=tmp_path)
    q, fqt, fq, fq2 = read_signal(tmp_path)
    assert q.shape == (4, 3)
    for i in range(4):
        r = oracle_mod.compute("all", frames, b, subs[i], nthreads=8)
        assert shell_error(fqt[i], r[0])[0] < 1e-5
        assert abs(fq[i] - r[1]) < 1e-5 * abs(r[0][0])
